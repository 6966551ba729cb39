"""Shared helpers of the GPU tests."""
import sys

import pytest


def announce(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def rel_close(got, want, rtol, floor=1.0):
    """|got - want| <= rtol * max(|want|, floor), elementwise; returns the worst ratio."""
    worst = 0.0
    for g, w in zip(got, want):
        worst = max(worst, abs(g - w) / (rtol * max(abs(w), floor)))
    return worst
