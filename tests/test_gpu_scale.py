"""The benchmark's own sizes under parity checks (the reference's check at
main.c:193-258 is the 1-norm printout of Q^T Q - I and QR - A; these are
size-independent restatements of it on the device):

* 16384 x 16384, nb = 256, ib = 32 (the headline factorisation): the column
  norms of R equal those of A (Q orthogonal), and Q^T applied by the engine's
  own ormqr to 8 sampled columns of A reproduces R's columns and zeros below
  -- a check that covers the reflectors V, TAU and T, not only R;
* 65536 x 4096 (the tall configuration): Q^T A(:, 0:nb) = [R(0:nb, 0:nb); 0]
  through dormqr, dormqr on a 65536 x 64 block equal to the same columns of
  that product, and the dorgqr Q^T array with orthonormal columns.
FP64, 1e-12 relative to the column's norm."""
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hq():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import hybrid3_amd
    hybrid3_amd.lib()
    return hybrid3_amd


def colnorm_rel(torch, pristine, A, MT, NT, nb):
    a = pristine.view(NT, MT, nb, nb)   # [tj][ti][col][row]
    f = A.view(NT, MT, nb, nb)
    tj = torch.arange(NT, device="cuda").view(NT, 1, 1, 1)
    ti = torch.arange(MT, device="cuda").view(1, MT, 1, 1)
    col = torch.arange(nb, device="cuda").view(1, 1, nb, 1)
    row = torch.arange(nb, device="cuda").view(1, 1, 1, nb)
    upper = (ti * nb + row) <= (tj * nb + col)
    an = (a * a).sum(dim=(1, 3)).sqrt()
    rn = (f * f * upper).sum(dim=(1, 3)).sqrt()
    return ((an - rn).abs() / an).max().item()


def test_headline_16384_reflectors_reproduce_R(hq):
    import torch
    n, nb, ib = 16384, 256, 32
    MT = NT = n // nb
    print("16384^2 nb=256 ib=32", flush=True)
    pristine = torch.empty(MT * NT * nb * nb, dtype=torch.float64, device="cuda")
    hq.fill_uniform_dev(MT, NT, nb, 20151130, pristine)
    A = pristine.clone()
    TAU = torch.zeros(MT * NT * nb, dtype=torch.float64, device="cuda")
    T = torch.zeros(MT * NT * ib * nb, dtype=torch.float64, device="cuda")
    hq.tile_qr_dev(MT, NT, nb, ib, A, TAU, T)
    hq.sync()
    assert colnorm_rel(torch, pristine, A, MT, NT, nb) < 1e-12
    # 8 sampled global columns of A into one tile column of X, Q^T X by ormqr
    cols = [0, 1, 255, 256, 4097, 9000, 16000, 16383]
    P = pristine.view(NT, MT, nb, nb)   # [tj][ti][col][row]
    F = A.view(NT, MT, nb, nb)
    X = torch.zeros(MT, nb, nb, dtype=torch.float64, device="cuda")   # [ti][col][row]
    for c, j in enumerate(cols):
        X[:, c, :] = P[j // nb, :, j % nb, :]
    hq.tile_ormqr_dev(MT, NT, nb, ib, A, TAU, X.view(-1), 1, T=T)
    hq.sync()
    for c, j in enumerate(cols):
        got = X[:, c, :].reshape(-1)                   # Q^T a_j, all rows
        rcol = F[j // nb, :, j % nb, :].reshape(-1).clone()
        rows = torch.arange(MT * nb, device="cuda")
        rcol[rows > j] = 0.0                           # R's column j: rows <= j
        scale = torch.linalg.norm(P[j // nb, :, j % nb, :]).item()
        err = (got - rcol).abs().max().item() / scale
        assert err < 1e-12, (j, err)


def test_tall_65536x4096_orgqr_ormqr(hq):
    import torch
    m, n, nb, ib = 65536, 4096, 256, 32
    MT, NT = m // nb, n // nb
    print("65536x4096 nb=256 ib=32", flush=True)
    pristine = torch.empty(MT * NT * nb * nb, dtype=torch.float64, device="cuda")
    hq.fill_uniform_dev(MT, NT, nb, 20151130, pristine)
    A = pristine.clone()
    TAU = torch.zeros(MT * NT * nb, dtype=torch.float64, device="cuda")
    T = torch.zeros(MT * NT * ib * nb, dtype=torch.float64, device="cuda")
    hq.tile_qr_dev(MT, NT, nb, ib, A, TAU, T)
    hq.sync()
    assert colnorm_rel(torch, pristine, A, MT, NT, nb) < 1e-12
    # dormqr on tile column 0 of A: Q^T A(:, 0:nb) = [R(0:nb, 0:nb); 0]
    C = pristine[: MT * nb * nb].clone()
    hq.tile_ormqr_dev(MT, NT, nb, ib, A, TAU, C, 1, T=T)
    hq.sync()
    qc = C.view(MT, nb, nb)           # [tile row][col][row]
    r00 = A[: nb * nb].view(nb, nb)   # [col][row]
    upper = torch.tril(torch.ones(nb, nb, dtype=torch.bool, device="cuda"))   # row <= col
    rmax = r00.abs().max().item()
    assert ((qc[0] - r00) * upper).abs().max().item() / rmax < 1e-12
    below = torch.cat([(qc[0] * ~upper).reshape(-1), qc[1:].reshape(-1)]).abs().max().item() / rmax
    assert below < 1e-12
    # the 65536 x 64 block of the benchmark: the first 64 columns, same result
    C64 = torch.zeros(MT * nb * nb, dtype=torch.float64, device="cuda").view(MT, nb, nb)
    C64[:, :64, :] = pristine[: MT * nb * nb].view(MT, nb, nb)[:, :64, :]
    hq.tile_ormqr_dev(MT, NT, nb, ib, A, TAU, C64.view(-1), 1, T=T)
    hq.sync()
    assert (C64[:, :64, :] - qc[:, :64, :]).abs().max().item() / rmax < 1e-12
    # the same block through the narrow-C entry (16-column tasks): the same
    # columns to rounding, the zero columns untouched
    C64n = torch.zeros(MT * nb * nb, dtype=torch.float64, device="cuda").view(MT, nb, nb)
    C64n[:, :64, :] = pristine[: MT * nb * nb].view(MT, nb, nb)[:, :64, :]
    hq.tile_ormqr_cols_dev(MT, NT, nb, ib, A, TAU, C64n.view(-1), 1, 64, T=T)
    hq.sync()
    assert (C64n[:, :64, :] - qc[:, :64, :]).abs().max().item() / rmax < 1e-12
    assert C64n[:, 64:, :].abs().max().item() == 0.0
    del C64n
    # dorgqr: the Q array (Q^T applied to the first n columns of I) has
    # orthonormal columns; a sample of 64 of them
    Q = torch.empty_like(pristine)
    hq.tile_orgqr_dev(MT, NT, nb, ib, A, TAU, Q, T=T)
    hq.sync()
    Qv = Q.view(NT, MT, nb, nb)
    S = torch.cat([Qv[j // nb, :, j % nb, :].reshape(-1, 1) for j in range(0, n, 64)], dim=1)   # m x 64
    G = S.t() @ S
    assert (G - torch.eye(G.shape[0], dtype=torch.float64, device="cuda")).abs().max().item() < 1e-12
