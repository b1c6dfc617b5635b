"""Host side of the Lanczos solver (no GPU): argument validation happens
before any device call, and the closed form its GPU test relies on is pinned
against dense diagonalisation here."""
import numpy as np
import pytest

from desv import lanczos, workloads
from oracle import sv_oracle as O


def test_tfim_free_fermion_closed_form_matches_dense():
    for n, seed in ((6, 1), (8, 2)):
        terms, J, h = workloads.tfim_terms(n, seed=seed)
        dense = np.linalg.eigvalsh(O.pauli_matrix(terms, n))[0]
        closed = -np.sum(np.linalg.svd(np.diag(h) + np.diag(J, 1), compute_uv=False))
        assert abs(dense - closed) < 1e-12


@pytest.mark.parametrize("kw,match", [
    (dict(which="XX"), "which"),
    (dict(k=0), "k must"),
    (dict(k=16), "k must"),
    (dict(ncv=65), "ncv"),
    (dict(k=3, ncv=3), "ncv"),
])
def test_argument_validation(kw, match):
    terms = [(1.0, {0: 'Z'}), (0.5, {1: 'X'}), (0.2, {3: 'Y'})]
    with pytest.raises(ValueError, match=match):
        lanczos.eigsh(terms, n_qubits=4, **kw)
    with pytest.raises(ValueError, match="n_qubits"):
        lanczos.eigsh(terms)
