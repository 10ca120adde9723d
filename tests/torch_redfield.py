"""Torch restatement of the Redfield block for the tests: autograd through
it (torch.linalg.eigh, C(w), Redfield tensor, basis change) is the reference
that the hand-written reverse mode (qsx_redfield_block_vjp) is checked
against.  Follows qspectra/dynamics/redfield.py:9-104 and bath.py:70-102;
its own values are checked against the oracle in test_gradients_host.py."""
import numpy as np
import torch

from qspectra_amd.constants import CM_FS


def debye_correlation(x, lam, gamma, T, real=False, matsubara_cutoff=1000):
    """Debye bath correlation C(x) on a real tensor x (bath.py:70-102): the
    complex form with the reference's truncated Matsubara sum (the k = 0
    term is zero) or, for real=True, (1/expm1(x/T) + 1) J(x) with
    J(x) = 2 lam gamma x / (gamma^2 + x^2); C(0) = lam (2T/gamma - i) or
    2 lam T / gamma.  Differentiable in x and lam."""
    zero = x == 0
    xs = torch.where(zero, torch.ones_like(x), x)
    if real:
        J = 2 * lam * gamma * xs / (gamma ** 2 + xs ** 2)
        val = (1 / torch.expm1(xs / T) + 1) * J
        c0 = 2 * lam * T / gamma
        return torch.where(zero, c0 * torch.ones_like(x), val).to(torch.complex128)
    nu = 2 * np.pi * T * torch.arange(1, matsubara_cutoff, dtype=torch.float64, device=x.device)
    a = nu / (nu ** 2 - gamma ** 2)
    xc = xs.to(torch.complex128)[..., None]
    msum = (a / (nu - 1j * xc)).sum(-1)
    cot = 1.0 / np.tan(gamma / (2 * T))
    val = lam * gamma * ((cot - 1j) / (gamma - 1j * xs.to(torch.complex128)) + 4 * T * msum)
    c0 = lam * (2 * T / gamma - 1j)
    return torch.where(zero, c0 * torch.ones_like(val), val)


def _block_eig(H1, X, rw):
    """(E rotating frame [B, nX], U [B, nX, nX], number operators' K [B, N, nX, nX])
    for Hilbert block X ('g' or 'e') of the one-exciton Hamiltonians H1."""
    B, N, _ = H1.shape
    if X == 'g':
        z = torch.zeros((B, 1), dtype=torch.float64, device=H1.device)
        return z, torch.ones((B, 1, 1), dtype=torch.float64, device=H1.device), \
            torch.zeros((B, N, 1, 1), dtype=torch.float64, device=H1.device)
    E, U = torch.linalg.eigh(H1)
    K = U[:, :, None, :] * U[:, :, :, None]          # K_i[a][c] = U[i,a] U[i,c]
    return E - rw, U, K


def redfield_block(H1, lam, gamma, T, rw, X, Y, secular=False, discard_imag_corr=False,
                   unit_convert=CM_FS):
    """Redfield Liouvillian block XY (X, Y in 'g', 'e') in the site basis for
    one-exciton Hamiltonians H1 [B, N, N] (lab frame, float64 tensor):
    L[(a + nX b), (c + nX d)] of -i[H, .] - R (redfield.py:9-104,
    liouville_space.py:68-131), times unit_convert.  Differentiable in H1
    and lam (torch autograd)."""
    EX, UX, KX = _block_eig(H1, X, rw)
    EY, UY, KY = (EX, UX, KX) if Y == X else _block_eig(H1, Y, rw)

    def corr(E):
        return debye_correlation(E[:, :, None] - E[:, None, :], lam, gamma, T,
                                 real=discard_imag_corr)           # C[c][a] = C(E_c - E_a)

    def gs(K, C):
        # Gs[a][c] = sum_i sum_b K_i[a][b] K_i[b][c] C[c][b]
        Kc = K.to(torch.complex128)
        return torch.einsum('niab,nibc,ncb->nac', Kc, Kc, C)

    CX = corr(EX)
    CY = CX if Y == X else corr(EY)
    GX, GY = gs(KX, CX), (gs(KX, CX) if Y == X else gs(KY, CY))
    nX, nY = EX.shape[1], EY.shape[1]
    dev = H1.device
    IX = torch.eye(nX, dtype=torch.complex128, device=dev)
    IY = torch.eye(nY, dtype=torch.complex128, device=dev)
    KXc, KYc = KX.to(torch.complex128), KY.to(torch.complex128)
    R = (torch.einsum('bd,nac->nabcd', IY, GX)
         + torch.einsum('ac,nbd->nabcd', IX, GY.conj())
         - torch.einsum('niac,nibd,ndb->nabcd', KXc, KYc, CY.conj())
         - torch.einsum('niac,nibd,nca->nabcd', KXc, KYc, CX))
    w = (EX[:, :, None] - EY[:, None, :]).to(torch.complex128)       # [n, a, b]
    L = -1j * torch.einsum('nab,ac,bd->nabcd', w, IX, IY) - R
    if secular:
        ia = torch.arange(nX, device=dev)
        ib = torch.arange(nY, device=dev)
        keep = (ia[:, None, None, None] == ia[None, None, :, None]) & \
               (ib[None, :, None, None] == ib[None, None, None, :])
        if X == Y:
            keep = keep | ((ia[:, None, None, None] == ib[None, :, None, None])
                           & (ia[None, None, :, None] == ib[None, None, None, :]))
        L = L * keep
    UXc, UYc = UX.to(torch.complex128), UY.to(torch.complex128)
    L = torch.einsum('npa,nqb,nabcd,nrc,nsd->npqrs', UXc, UYc, L, UXc, UYc)
    B = L.shape[0]
    L = L.permute(0, 2, 1, 4, 3).reshape(B, nX * nY, nX * nY)
    return L * unit_convert
