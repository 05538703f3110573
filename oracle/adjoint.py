"""Derivatives of the grid exchange-correlation energy, restated in numpy.

TEST INFRASTRUCTURE ONLY: imported by tests/, never by the product package.

The reference has no adjoint.  Its GTO gradient (src/cgf.cpp:67-94, 164-172)
omits the primitive norm and applies the contraction coefficient twice, so it
is not followed.  What is differentiated is the reference's energy

    E_xc = sum_p w_p e(rho_p),   rho_p = c rho_raw_p,   rho_raw_p = 2 phi_p^T P phi_p,
    c = Z_tot / S,   S = sum_p w_p rho_raw_p

(renormalisation src/moleculargrid.cpp:132-146, E_xc src/dft.cpp:394-433),
with v = de/drho the reference's own potential (its exactness is pinned by
finite differences in the LDA pins).  The chain rule gives, with
Lambda = sum_p w_p v_p rho_p and k = Lambda / Z_tot (k = 0 when c is held):

    dE/dP    = 2 c Phi^T diag(w (v - k)) Phi
    dE/dR_A  = sum_p (dw_p/dR_A) (e_p - k rho_p)
             + 4 sum_p w_p c (v_p - k) sum_i T_pi dphi_i(r_p)/dR_A,   T = Phi P,

where dphi_i/dR_A = -grad phi_i for functions centred on A, plus grad phi_i
for every i when the point belongs to A's cell (the grid moves with its
atom), and dw_p/dR_A is the Becke-weight derivative (the point of cell a
moves with a, so dwb/dR_a = -sum_{B != a} dwb/dR_B).  Every function here is
pinned by central finite differences of the oracle's own E_xc (and S) in
tests/test_oracle_adjoint.py.
"""
import numpy as np

from oracle import oracle as O

COMPONENTS = {0: [(0, 0, 0)], 1: [(1, 0, 0), (0, 1, 0), (0, 0, 1)],
              2: [(2, 0, 0), (1, 1, 0), (1, 0, 1), (0, 2, 0), (0, 1, 1), (0, 0, 2)]}


def _pow(x, k):
    return np.ones_like(x) if k == 0 else x ** k


def basis_values(mol, xyz):
    """phi [npts, nbf], grad phi [npts, nbf, 3] and the centre atom of each CGF."""
    s = mol.shells()
    phis, grads, centre = [], [], []
    for k in range(len(s["L"])):
        a = int(s["atom"][k])
        d = xyz - mol.xyz[a]
        r2 = (d * d).sum(1)
        off, n = int(s["offset"][k]), int(s["nprim"][k])
        al, co = s["alpha"][off:off + n], s["coef"][off:off + n]
        e = np.exp(-np.outer(r2, al))
        for (l, m, q) in COMPONENTS[int(s["L"][k])]:
            cn = co * np.array([O.gto_norm(x, l, m, q) for x in al])
            R = e @ cn
            Rp = e @ (cn * -2.0 * al)  # (1/x) d/dx of the radial part
            px, py, pz = _pow(d[:, 0], l), _pow(d[:, 1], m), _pow(d[:, 2], q)
            poly = px * py * pz
            dpx = l * _pow(d[:, 0], l - 1) if l else np.zeros_like(R)
            dpy = m * _pow(d[:, 1], m - 1) if m else np.zeros_like(R)
            dpz = q * _pow(d[:, 2], q - 1) if q else np.zeros_like(R)
            phis.append(R * poly)
            grads.append(np.stack([dpx * py * pz * R + d[:, 0] * poly * Rp,
                                   px * dpy * pz * R + d[:, 1] * poly * Rp,
                                   px * py * dpz * R + d[:, 2] * poly * Rp], axis=1))
            centre.append(a)
    return np.stack(phis, 1), np.stack(grads, 1), np.asarray(centre)


def becke_cutoff_and_ratio(mu):
    """s(mu) = (1 - f(f(f(mu))))/2, f(x) = 1.5x - 0.5x^3, and s'(mu)/s(mu) (0 where s = 0)."""
    f1 = 1.5 * mu - 0.5 * mu ** 3
    f2 = 1.5 * f1 - 0.5 * f1 ** 3
    f3 = 1.5 * f2 - 0.5 * f2 ** 3
    s = 0.5 * (1.0 - f3)
    sp = -0.5 * (1.5 - 1.5 * f2 ** 2) * (1.5 - 1.5 * f1 ** 2) * (1.5 - 1.5 * mu ** 2)
    ratio = np.divide(sp, s, out=np.zeros_like(s), where=s > 0)
    return s, ratio


def becke_weight_gradient(mol, xyz, owner):
    """wb [npts] and d wb / d R_B [npts, natoms, 3] (points move with their owner atom)."""
    R = mol.xyz
    na = len(R)
    diff = xyz[:, None, :] - R[None]                 # r - R_n
    dist = np.linalg.norm(diff, axis=2)
    u = diff / dist[:, :, None]
    Rnm = np.linalg.norm(R[:, None] - R[None], axis=2) + np.eye(na)
    e = (R[:, None] - R[None]) / Rnm[:, :, None]     # e_nm = (R_n - R_m)/R_nm
    mu = (dist[:, :, None] - dist[:, None, :]) / Rnm[None]
    s, ratio = becke_cutoff_and_ratio(mu)
    eye = np.eye(na, dtype=bool)[None]
    s = np.where(eye, 1.0, s)
    ratio = np.where(eye, 0.0, ratio)
    Pn = s.prod(2)                                    # [p, n]
    Zs = Pn.sum(1)
    idx = np.arange(len(xyz))
    wb = Pn[idx, owner] / Zs
    # dP_n/dR_B, r fixed: [p, n, B, 3]
    # n == B: P_B sum_m ratio(mu_Bm) * (-(u_B + mu_Bm e_Bm)/R_Bm)
    # n != B: P_n ratio(mu_nB) (u_B + mu_nB e_nB)/R_nB
    t = (u[:, None, :, :] + mu[:, :, :, None] * e[None]) / Rnm[None, :, :, None]  # [p, n, m, 3]
    dP = Pn[:, :, None, None] * ratio[:, :, :, None] * t                           # n != B part
    self_term = -(Pn[:, :, None] * (ratio[:, :, :, None] * (u[:, :, None, :] + mu[:, :, :, None] * e[None])
                                    / Rnm[None, :, :, None]).sum(2))                # [p, n, 3]
    dP[:, np.arange(na), np.arange(na), :] = self_term
    dwb = (dP[idx, owner] - wb[:, None, None] * dP.sum(1)) / Zs[:, None, None]      # [p, B, 3]
    own = -(np.where(np.arange(na)[None, :, None] == owner[:, None, None], 0.0, dwb)).sum(1)
    dwb[idx, owner] = own
    return wb, dwb


def xc_adjoint(mol, nrad, nang, lmax, P, hold_c=False):
    """(E_xc, dE/dP [nbf, nbf], dE/dR [natoms, 3]) of the oracle's E_xc at density matrix P."""
    g = O.Grid(mol, nrad, nang, lmax)
    g.set_density(P)
    g.correct_densities()
    _, exc = g.xc()
    xyz, w, wb = g.points()
    rho = g.densities()
    e, wv = g.xc_parts()
    phi, dphi, centre = basis_values(mol, xyz)
    rho_raw = 2.0 * np.einsum("pi,ij,pj->p", phi, P, phi)
    Ztot = float(np.sum(mol.Z))
    c = Ztot / float(w @ rho_raw)
    k = 0.0 if hold_c else float(wv @ rho) / Ztot
    ub = c * (wv - k * w)
    dEdP = 2.0 * phi.T @ (ub[:, None] * phi)
    T = phi @ P
    G = 4.0 * ub[:, None, None] * T[:, :, None] * dphi          # [p, i, 3]
    na = mol.natoms
    cell = nrad * nang
    owner = np.arange(len(w)) // cell
    grad = np.zeros((na, 3))
    Gi = G.sum(0)
    for i, a in enumerate(centre):
        grad[a] -= Gi[i]
    Gp = G.sum(1)
    for a in range(na):
        grad[a] += Gp[owner == a].sum(0)
    wb2, dwb = becke_weight_gradient(mol, xyz, owner)
    w0 = np.divide(w, wb, out=np.zeros_like(w), where=wb > 0)
    grad += np.einsum("p,pbx->bx", w0 * (e - k * rho), dwb)
    return exc, dEdP, grad


def exc_at(mol, nrad, nang, lmax, P):
    """E_xc and S = sum w rho_raw of the oracle pipeline at density matrix P."""
    g = O.Grid(mol, nrad, nang, lmax)
    g.set_density(P)
    S = g.calculate_density()
    g.correct_densities()
    _, exc = g.xc()
    return exc, S
