"""A float64 torch restatement of the chain from the colour-charge normals to
dE/deta, differentiated by torch autograd: the independent reference
gradient for the library's hand-written reverse mode (test infrastructure
only; CPU tensors).

Forward, as the reference computes it:
  rho^a_k(x)  = g sqrt(g2mu2(x)/Ny) z^a_k(x)                 Init.cpp:1697-1705
  A^a_k       = FFT^-1[ FFT[rho^a_k] / (k~^2 + m~^2) ]          Init.cpp:1707-1747
                k~^2 = 4 (sin^2(k_x/2) + sin^2(k_y/2)); zero mode dropped at m = 0
  V           = prod_k exp(-i A_k^a t^a), later slices on the left   1750-1768
  U^(n)_i(x)  = V_n(x) V_n(x+i)^+ (clamped neighbours)            2459-2491
  U_i(x)      : F_a(U) = 0, U = exp(i alpha^a t^a) U^(1) U^(2)   3344-3481
  pi          = -(2i/g) A^eta from the four links at x           2519-2690
  leapfrog    evolvePhi/evolveU kick at tau = 0, then itmax-1 steps of
              evolvePi, evolveE, evolvePhi, evolveU       Evolution.cpp:35-431
  dE/deta     = tau a^2 hbarc sum_x T^tautau at itmax            659-926
with torch.linalg.matrix_exp for every exponential (not the Cayley-Hamilton
closed form), torch.fft for the transforms (not the LDS Stockham code),
Newton with an autograd Jacobian for the root (iterated to round-off), and
implicit differentiation of the root (the Jacobian again from autograd).
"""
import numpy as np
import torch

HBARC = 0.1973269718
DT = torch.float64
CT = torch.complex128


def generators():
    t = torch.zeros(8, 3, 3, dtype=CT)
    t[0, 0, 1] = t[0, 1, 0] = 0.5
    t[1, 0, 1], t[1, 1, 0] = -0.5j, 0.5j
    t[2, 0, 0], t[2, 1, 1] = 0.5, -0.5
    t[3, 0, 2] = t[3, 2, 0] = 0.5
    t[4, 0, 2], t[4, 2, 0] = -0.5j, 0.5j
    t[5, 1, 2] = t[5, 2, 1] = 0.5
    t[6, 1, 2], t[6, 2, 1] = -0.5j, 0.5j
    t[7] = torch.diag(torch.tensor([1.0, 1.0, -2.0], dtype=CT)) / (2 * np.sqrt(3))
    return t


GEN = generators()


def dag(m):
    return m.conj().transpose(-1, -2)


def tr(m):
    return torch.diagonal(m, dim1=-2, dim2=-1).sum(-1)


def expi(q):
    """exp(i q^a t^a) for real q [..., 8]"""
    H = torch.einsum("...a,aij->...ij", q.to(CT), GEN)
    return torch.linalg.matrix_exp(1j * H)


def shift(f, axis, d):
    """f at the clamped neighbour (i + d along axis) of every site; f is
    [N, N, ...] with axis 0 = x (i), axis 1 = y (j)"""
    N = f.shape[axis]
    idx = torch.clamp(torch.arange(N) + d, 0, N - 1)
    return torch.index_select(f, axis, idx)


# ---- Wilson lines ----------------------------------------------------------------

def wilson_lines(z, g2mu2, g, m_lat, Ny):
    """z [2, Ny, N, N, 8] normals, g2mu2 [2, N, N] -> V [2, N, N, 3, 3]"""
    N = z.shape[2]
    q = torch.arange(N, dtype=DT) * 2 * np.pi / N
    s2 = 4 * torch.sin(q / 2) ** 2
    kt2 = s2[:, None] + s2[None, :]
    if m_lat == 0:
        mult = torch.where(kt2 != 0, 1.0 / torch.where(kt2 != 0, kt2, torch.ones_like(kt2)), torch.zeros_like(kt2))
    else:
        mult = 1.0 / (kt2 + m_lat * m_lat)
    Vs = []
    for n in range(2):
        amp = g * torch.sqrt(g2mu2[n] / Ny)
        V = torch.eye(3, dtype=CT).expand(N, N, 3, 3)
        for k in range(Ny):
            rho = amp[..., None] * z[n, k]  # [N, N, 8]
            R = torch.fft.fft2(rho.to(CT), dim=(0, 1))
            A = torch.fft.ifft2(R * mult[..., None], dim=(0, 1)).real
            V = expi(-A) @ V
        Vs.append(V)
    return torch.stack(Vs)


def links(V):
    """U_x(x) = V(x) V(x+x)^+, U_y likewise; [N, N, 3, 3] each"""
    return V @ dag(shift(V, 0, 1)), V @ dag(shift(V, 1, 1))


# ---- the tau = 0 root ------------------------------------------------------------

def residual(beta, U0, U1, U2):
    """F_a = -Im[tr t^a (W - W^+) + tr t^a (W U^+ - U W^+)] for
    U = exp(i beta t) U0, W = U1 + U2 (U0 a fixed point of the coset
    SU(3) U1 U2, so that U stays on it)"""
    W = U1 + U2
    U = expi(beta) @ U0
    M = W - dag(W) + W @ dag(U) - U @ dag(W)
    return -torch.einsum("aij,...ji->...a", GEN, M).imag


def _jacobian(beta, U0, U1, U2):
    """dF/dbeta [n, 8, 8] by forward-mode autograd over the 8 directions"""
    cols = []
    for b in range(8):
        e = torch.zeros_like(beta)
        e[:, b] = 1.0
        _, jv = torch.autograd.functional.jvp(lambda x: residual(x, U0, U1, U2), (beta,), (e,))
        cols.append(jv)
    return torch.stack(cols, -1)


def solve_root(U0, U1, U2, iters=60, tol=1e-14):
    """Newton on beta (no graph), from beta = 0 (U = U0)"""
    with torch.no_grad():
        b = torch.zeros(U1.shape[0], 8, dtype=DT)
        for _ in range(iters):
            F = residual(b, U0, U1, U2)
            if F.abs().sum(-1).max() < tol:
                break
            J = _jacobian(b, U0, U1, U2)
            b = b - torch.linalg.solve(J, F)
        return b


class RootBeta(torch.autograd.Function):
    """beta(U1, U2) with F(beta; U1, U2) = 0 for U = exp(i beta t) U0, U0
    held fixed; backward by the implicit function theorem:
    lambda = J^-T g, dL/dU_n = -(lambda . dF/dU_n)"""

    @staticmethod
    def forward(ctx, U0, U1, U2):
        b = solve_root(U0, U1, U2)
        ctx.save_for_backward(b, U0, U1, U2)
        return b

    @staticmethod
    def backward(ctx, gb):
        b, U0, U1, U2 = ctx.saved_tensors
        with torch.enable_grad():
            J = _jacobian(b, U0, U1, U2)
            lam = torch.linalg.solve(J.transpose(-1, -2), gb)
            u1 = U1.detach().requires_grad_(True)
            u2 = U2.detach().requires_grad_(True)
            F = residual(b.detach(), U0, u1, u2)
            g1, g2 = torch.autograd.grad(F, (u1, u2), grad_outputs=-lam)
        return None, g1, g2


def tau0_fields(VA, VB, g, start=None):
    """links, root, pi = -(2i/g) A^eta (Init.cpp:2437-2690).  The root is
    sought near `start` (Ux, Uy [N, N, 3, 3], e.g. another solver's roots,
    which picks the same root where several exist), else near U1 U2."""
    N = VA.shape[0]
    U1x, U1y = links(VA)
    U2x, U2y = links(VB)
    out = []
    for k, (A1, A2) in enumerate(((U1x, U2x), (U1y, U2y))):
        a1, a2 = A1.reshape(-1, 3, 3), A2.reshape(-1, 3, 3)
        U0 = (a1 @ a2).detach() if start is None else torch.as_tensor(start[k], dtype=CT).reshape(-1, 3, 3)
        b = RootBeta.apply(U0, a1, a2)
        out.append((expi(b) @ U0).reshape(N, N, 3, 3))
    Ux, Uy = out
    acc = torch.zeros(N, N, 3, 3, dtype=CT)
    for ax, (A1, A2, U) in enumerate(((U1x, U2x, Ux), (U1y, U2y, Uy))):
        D = A1 - A2
        Dd = dag(A1) - dag(A2)
        acc = acc + (D @ dag(U) - D - U @ Dd + Dd)
        Dm, Ddm, Um = shift(D, ax, -1), shift(Dd, ax, -1), shift(U, ax, -1)
        acc = acc + (-dag(Um) @ Dm + Dm + Ddm @ Um - Ddm)
    pi = (-2j / g) * (acc / 8)
    return Ux, Uy, pi


# ---- leapfrog and T^tautau -------------------------------------------------------

def traceless(m):
    return m - tr(m)[..., None, None] / 3 * torch.eye(3, dtype=CT)


def step(st, g, dtau, tau, half=False):
    """one Evolution::run step at time tau (Evolution.cpp:35-293)"""
    Ux, Uy, E1, E2, phi, pi = st
    dt = dtau / 2 if half else dtau
    # evolvePi
    pX = Ux @ shift(phi, 0, 1) @ dag(Ux)
    pY = Uy @ shift(phi, 1, 1) @ dag(Uy)
    uxm, uym = shift(Ux, 0, -1), shift(Uy, 1, -1)
    pmX = dag(uxm) @ shift(phi, 0, -1) @ uxm
    pmY = dag(uym) @ shift(phi, 1, -1) @ uym
    pi_n = pi + (dt / tau) * (pX + pmX + pY + pmY - 4 * phi)
    # evolveE: plaquettes U_12, U_1,-2, U_2,-1 and the phi commutators
    U12 = Ux @ shift(Uy, 0, 1) @ dag(shift(Ux, 1, 1)) @ dag(Uy)
    Uy_pxmy = shift(shift(Uy, 0, 1), 1, -1)
    Ux_my, Uy_my = shift(Ux, 1, -1), shift(Uy, 1, -1)
    U1m2 = Ux @ dag(Uy_pxmy) @ dag(Ux_my) @ Uy_my
    Ux_mxpy = shift(shift(Ux, 0, -1), 1, 1)
    Uy_mx, Ux_mx = shift(Uy, 0, -1), shift(Ux, 0, -1)
    U2m1 = Uy @ dag(Ux_mxpy) @ dag(Uy_mx) @ Ux_mx
    c1 = 1j * tau * dt / (2 * g * g)
    c2 = 1j * dt / tau
    t1 = traceless(U12 + U1m2 - dag(U12) - dag(U1m2))
    E1_n = traceless(E1 + c1 * t1 + c2 * (pX @ phi - phi @ pX))
    t1 = traceless(dag(U12) + U2m1 - U12 - dag(U2m1))
    E2_n = traceless(E2 + c1 * t1 + c2 * (pY @ phi - phi @ pY))
    if half:
        return Ux, Uy, E1_n, E2_n, phi, pi_n
    # evolvePhi, evolveU
    phi_n = phi + (tau + dtau / 2) * dtau * pi_n
    c = 1j * g * g * dtau / (tau + dtau / 2)
    eye = torch.eye(3, dtype=CT)

    def upd(E, U):
        X = c * E
        return (eye + X @ (eye + X / 2)) @ U

    return upd(E1_n, Ux), upd(E2_n, Uy), E1_n, E2_n, phi_n, pi_n


def ttautau(st, g, tau):
    """T^tautau in lattice units (Evolution.cpp:754-905)"""
    Ux, Uy, E1, E2, phi, pi = st
    N = Ux.shape[0]
    e1p, e2p = shift(E1, 1, 1), shift(E2, 0, 1)
    el = (g * g / tau / tau) * (tr(E1 @ E1) + tr(e1p @ e1p) + tr(E2 @ E2) + tr(e2p @ e2p)).real / 2
    pX, pY, pXY = shift(pi, 0, 1), shift(pi, 1, 1), shift(shift(pi, 0, 1), 1, 1)
    lng = (tr(pi @ pi) + tr(pX @ pX) + tr(pY @ pY) + tr(pXY @ pXY)).real / 4
    plaq = Ux @ shift(Uy, 0, 1) @ dag(shift(Ux, 1, 1)) @ dag(Uy)
    rtr = tr(plaq).real
    last = (torch.arange(N) == N - 1)[:, None].expand(N, N)
    rtr = torch.where(last, torch.full_like(rtr, 3.0), rtr)
    mag = 2.0 / g ** 2 * (3.0 - rtr)
    phX, phY = shift(phi, 0, 1), shift(phi, 1, 1)
    phXY = shift(phX, 1, 1)
    uxpY, uypX = shift(Ux, 1, 1), shift(Uy, 0, 1)

    def sq(d):
        return tr(d @ d).real

    A = sq(phi - Ux @ phX @ dag(Ux))
    B = sq(phY - uxpY @ phXY @ dag(uxpY))
    Cc = sq(phi - Uy @ phY @ dag(Uy))
    D = sq(phX - uypX @ phXY @ dag(uypX))
    return el + lng + mag + 0.5 / tau / tau * (A + B + Cc + D)


def dedeta(z, g2mu2, g, m_lat, Ny, L, dtau, itmax, start=None):
    """dE/deta [GeV] from the normals z [2, Ny, N, N, 8] (requires_grad);
    `start`: starting points of the root finder (tau0_fields)"""
    N = z.shape[2]
    a = L / N
    V = wilson_lines(z, g2mu2, g, m_lat, Ny)
    Ux, Uy, pi = tau0_fields(V[0], V[1], g, start)
    zero = torch.zeros_like(pi)
    # tau = 0 kick: phi <- phi + (dtau/2) dtau pi; U unchanged (E = 0)
    st = (Ux, Uy, zero, zero, (dtau / 2) * dtau * pi, pi)
    for it in range(1, itmax):
        st = step(st, g, dtau, it * dtau)
    tau = itmax * dtau
    T = ttautau(st, g, tau) / a ** 4
    tau_fm = itmax * dtau * a
    return tau_fm * a * a * HBARC * T.sum()
