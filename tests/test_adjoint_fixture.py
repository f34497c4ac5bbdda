"""The committed FP64 values of the differentiation workload
(tests/golden/adjoint_128.npz, written by tests/golden/make_adjoint_fixture.py)
on the CPU: the generator still produces the workload they were computed on
(SHA-256 of the flag and vx fields), and the fixture is self-consistent --
the FP64 host adjoint's dJ/du_x agrees with the central differences of the
FP64 forward run on the 32 checked cells (the two are independent
computations: one reverse sweep, 64 forward runs)."""
from __future__ import annotations

import hashlib
from pathlib import Path

import numpy as np

from culbm_amd import cases

FIXTURE = Path(__file__).resolve().parent / "golden" / "adjoint_128.npz"


def test_fixture_matches_the_generator():
    g = np.load(FIXTURE)
    st = cases.config_adjoint(128)[0]
    assert hashlib.sha256(np.ascontiguousarray(st.flag).tobytes()).hexdigest() == str(g["flag_sha256"])
    assert hashlib.sha256(np.ascontiguousarray(st.vx).tobytes()).hexdigest() == str(g["vx_sha256"])
    inlet = (st.flag & cases.EQU) != 0
    inlet[:, :, 1:] = False
    assert np.array_equal(np.argwhere(inlet), g["inlet_cells"])
    assert list(g["steps"]) == [300, 251, 300, 20] and float(g["h"]) == 1e-4


def test_host_adjoint_agrees_with_finite_differences():
    g = np.load(FIXTURE)
    cells = [tuple(c) for c in g["inlet_cells"]]
    index = {c: k for k, c in enumerate(cells)}
    adj = np.array([g["grad_ux_inlet"][index[tuple(c)]] for c in g["fd_cells"]])
    fd = g["fd_grad_ux"]
    assert fd.size == 32
    # central differences at h = 1e-4 carry an O(h^2) truncation error
    assert np.max(np.abs(adj - fd) / np.abs(fd)) <= 1e-6
