"""A PEPS checkpoint read back without the library's loader.

The state file ("HBPP" v1: magic, u32 version, precision, n_sites, n_edges,
int32 edge pairs, per site u32 order + int32 shape + complex128 column-major
data, per message int32 D + complex128 D x D) is parsed here by an
independent numpy reader and compared with the oracle's state after the same
circuit (the reference's restart stores the same V and E, iotps.h:37-105):
bond dims and the Lambda on every bond (messages are diag(Lambda) after
Truncation) to 1e-10, and <Z>, <ZZ>, <XY> evaluated by the oracle's Measure
on the file's tensors to 1e-10 of the oracle's values on its own state
(gauge-invariant: the file's sites may differ from the oracle's by bond
gauges)."""
import struct

import numpy as np
import pytest

from gpu_util import announce
from oracle import circuits as oc
from oracle import tnbp as ot

pytestmark = pytest.mark.gpu


def _read_hbpp(path):
    buf = open(path, "rb").read()
    assert buf[:4] == b"HBPP"
    ver, prec, ns, ne = struct.unpack_from("<4I", buf, 4)
    off = 20
    flat = struct.unpack_from(f"<{2 * ne}i", buf, off)
    off += 8 * ne
    edges = [(flat[2 * i], flat[2 * i + 1]) for i in range(ne)]
    sites = []
    for _ in range(ns):
        (order,) = struct.unpack_from("<I", buf, off)
        off += 4
        shape = struct.unpack_from(f"<{order}i", buf, off)
        off += 4 * order
        n = int(np.prod(shape))
        sites.append(np.frombuffer(buf, dtype="<c16", count=n, offset=off).reshape(shape, order="F"))
        off += 16 * n
    msgs = []
    for _ in range(2 * ne):
        (d,) = struct.unpack_from("<i", buf, off)
        off += 4
        msgs.append(np.frombuffer(buf, dtype="<c16", count=d * d, offset=off).reshape((d, d), order="F"))
        off += 16 * d * d
    assert off == len(buf)
    return ver, prec, edges, sites, msgs


def test_checkpoint_file_read_independently_matches_oracle(tmp_path):
    from tnbp_amd import circuits as pc
    from tnbp_amd.peps import AbsorbTPO, InitTensorProductState, Truncation, bp_loop, save_peps
    from tnbp_amd.tci import Context
    announce("checkpoint file vs oracle state")
    rng = np.random.default_rng(6)
    edges = pc.square_lattice(3, 3)
    layers = [pc.random_layer(rng, edges, 9) for _ in range(3)]
    with Context(0) as ctx:
        p = InitTensorProductState(ctx, edges)
        for g in layers:
            AbsorbTPO(ctx, pc.layer_tpo(edges, g), p)
            bp_loop(ctx, p, 200, 1e-13)
            Truncation(ctx, p, 4, 1e-10)
        path = tmp_path / "state.hbpp"
        save_peps(p, path)
        p.close()
    ver, prec, fedges, V, E = _read_hbpp(path)
    assert ver == 1 and prec == 0 and fedges == [tuple(e) for e in edges]
    st, _ = oc.run_tpo_layers(edges, [oc.layer_tpo(edges, g) for g in layers], 200, 1e-13, 4, 1e-10)
    ne = len(edges)
    for m in range(ne):
        assert E[m].shape == st.E[m].shape
        np.testing.assert_allclose(np.diag(E[m]).real, np.diag(st.E[m]).real, rtol=1e-10, atol=1e-12)
        np.testing.assert_allclose(E[m], np.diag(np.diag(E[m])), atol=0)  # diag(Lambda) exactly
    fst = ot.State(edges, [v.copy() for v in V], list(range(9)), [e.copy() for e in E], list(range(ne)))
    ops1 = [oc.gate_tensor(np.diag([1.0, -1.0]).astype(complex))] * 9
    np.testing.assert_allclose(ot.measure_sites(fst, list(range(9)), ops1),
                               ot.measure_sites(st, list(range(9)), ops1), atol=1e-10)
    X = np.array([[0, 1], [1, 0]], dtype=complex)
    Y = np.array([[0, -1j], [1j, 0]])
    for u in (np.diag([1.0, -1, -1, 1]).astype(complex), np.kron(Y, X)):
        o = [oc.gate_tensor(u)] * ne
        np.testing.assert_allclose(ot.measure_edges(fst, edges, o), ot.measure_edges(st, edges, o),
                                   atol=1e-10)
