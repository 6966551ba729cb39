"""The parity oracle: a CPU restatement (numpy) of the BP-PEPS path and the
recipe that builds the reference's own front-end code from source into
``oracle/_ref``.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py`` may import or run anything here, and
only as the checker; the product (``tnbp_amd``) never does.
"""
