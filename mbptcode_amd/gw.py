"""The space-time GW correlation self-energy on the GPU.

The reference's serial space-time route (GW/space_time.py:789-849) without its
mean-field and quasiparticle ends:

1. chi0 on the screening frequencies: proj(tau) on chi0's own time axis (one
   fused launch per time point), then the cosine transform onto the
   frequencies (space_time.py:607-631 through `polarizability_projected_sweep`).
   Both in FP64: the screening frequencies reach below the gap (rW starts at
   0.3 gap) while chi0's time axis is fitted on [gap, span], and the transform
   weights there reach 7e8 with cancelling signs (water-size instance, 16
   points), so FP32 proj(tau) would lose every digit;
2. Dyson per frequency, W - V on the grid (GW/space_time.py:382-390): with
   L = chol(V), W = L (I - L^T chi0 L)^-1 L^T by an FP64 Cholesky, a
   triangular solve and a symmetric product;
3. the cosine transform of W - V to the self-energy's time axis
   (GW/imaginary_time.py:181-195), Zt(tau) in FP32;
4. the branch sums Sigma^<_pp(tau), Sigma^>_pp(tau) of the requested states
   (`SigmaPairs.add`, imaginary_time.py:75-92), one fused kernel launch
   (`mbpt_chi0_sigma`) over every time point;
5. the cosine and sine transforms back to the output frequencies
   (imaginary_time.py:557-561):
   Sigma^c_pp(i w) = -1/2 [(S^> + S^<) C^T + i (S^> - S^<) S^T].

Everything is in the grid gauge (the auxiliary basis is the interpolation grid
and V = L L^T), so the screened interaction is formed on the grid once per
frequency and no per-time projection is needed.  The frequency axes, transform
weights and fit ranges are `grids.space_time_gw_grids`'.  Sigma^c(i w) is
returned on the Pade frequencies; `quasiparticle_energies_space_time` adds
the continuation and the quasiparticle equation (`qp`) with the static term
<Sigma_x - v_xc> from the caller (its build needs the mean field: pyscf).
"""
import time

import numpy as np
import torch

from . import grids
from .space_time import Chi0Builder, _dev, _device, factor_batch_size, solve_lower_blocked


def screened_correlation_grid(L, chi0, V=None, infos=None):
    """W - V on the grid for one chi0 (M, M) FP64: W = L A^-1 L^T with
    A = I - L^T chi0 L = C C^T, as (L C^-T)(L C^-T)^T, then minus V = L L^T
    (V itself when given, saving a product).  With `infos` (a list) the
    factorization's status tensor is appended for the caller to check, so
    that no host synchronisation happens here."""
    K = L.T @ (chi0 @ L)
    n = K.shape[0]
    A = torch.eye(n, dtype=K.dtype, device=K.device) - 0.5 * (K + K.T)
    C, info = torch.linalg.cholesky_ex(A)
    if infos is None:
        if int(info) != 0:
            raise np.linalg.LinAlgError('I - L^T chi0 L is not positive definite')
    else:
        infos.append(info)
    Y = solve_lower_blocked(C, L.T.contiguous())               # C^-1 L^T
    Wc = Y.T @ Y - (L @ L.T if V is None else V)
    return 0.5 * (Wc + Wc.T)


def screened_correlation_frequencies(builder, V, grid, products='fp32'):
    """(W - V)(i w_k) on the grid for every frequency of `grid`, (nfreq, M, M)
    FP64: proj(tau) on grid's time axis once (an FP64 builder), its cosine
    transform, Dyson.  The frequencies' I - L^T chi0 L are factorized in
    batches of `factor_batch_size` (rocSOLVER's batched Cholesky keeps the
    device busy where one factorization alone is latency-bound), and each
    W - V is written over its chi0.  Everything stays on the caller's stream
    and the factorizations' status is checked once, after all of them.  The
    solves are not run concurrently: on several streams, from one host
    thread or from one thread per stream, torch's rocBLAS / rocSOLVER calls
    corrupted some frequencies' results (1e-4 relative, at random).
    products 'fp32': the products K = L^T chi0 L in FP32 (library GEMMs,
    widened before I - K is formed); the Cholesky, the triangular solve and
    Y^T Y - V stay FP64: W - V is a difference of two matrices of V's size,
    and Y^T Y in FP32 moved Sigma by 3e-5 relative where FP32 K moves it by
    1e-8.  'fp64': every product in FP64."""
    if products not in ('fp32', 'fp64'):
        raise ValueError(f"products={products!r}: 'fp32' or 'fp64'")
    if builder.precision != 'fp64':
        raise ValueError('the screening frequencies need FP64 proj(tau): the '
                         'cosine weights onto them cancel over many digits')
    dev = builder.device
    Vd = _dev(V, dev)
    L = torch.linalg.cholesky(Vd)
    proj = torch.stack([builder.build([t], [1.0]) for t in grid.tau_points])
    C = torch.as_tensor(np.asarray(grid.cosft_wt, float), device=dev)
    chi = torch.tensordot(C, proj, dims=([1], [0]))
    del proj
    nf, n = chi.shape[0], chi.shape[-1]
    g = factor_batch_size(n, nf)
    eye = torch.eye(n, dtype=chi.dtype, device=dev)
    Lt = L.T.contiguous()
    A = torch.empty((g, n, n), dtype=chi.dtype, device=dev)
    if products == 'fp32':
        L32, Lt32 = L.float(), Lt.float()
        c32, Y32, K32 = (torch.empty((n, n), dtype=torch.float32, device=dev)
                         for _ in range(3))
    infos = []
    for k0 in range(0, nf, g):
        ks = range(k0, min(nf, k0 + g))
        for j, k in enumerate(ks):
            if products == 'fp32':
                c32.copy_(chi[k])
                torch.matmul(c32, L32, out=Y32)
                torch.matmul(Lt32, Y32, out=K32)
                K = K32.double()
            else:
                K = L.T @ (chi[k] @ L)
            torch.sub(eye, 0.5 * (K + K.T), out=A[j])
        C, info = torch.linalg.cholesky_ex(A[:len(ks)])
        infos.append(info)
        for j, k in enumerate(ks):
            Y = solve_lower_blocked(C[j], Lt)                  # C^-1 L^T
            Wc = Y.T @ Y - Vd
            chi[k] = 0.5 * (Wc + Wc.T)                         # W - V over chi0
        del C
    if int(torch.cat(infos).abs().sum()) != 0:
        raise np.linalg.LinAlgError('I - L^T chi0 L is not positive definite')
    return chi


def screened_interaction_tau(Wc, Ctw):
    """Zt(tau_t) = sum_k Ctw[t, k] (W - V)(i w_k), (ntau, M, M) FP32
    (imaginary_time.py:181-195)."""
    C = torch.as_tensor(np.asarray(Ctw, float), device=Wc.device)
    return torch.tensordot(C, Wc, dims=([1], [0])).to(torch.float32).contiguous()


def self_energy_transforms(axes, omega_out=None):
    """(Ctw, C, S): W(i w) -> tau cosine weights on rW, and the tau -> omega
    cosine and sine weights on rS (time_frequency.py:430-449)."""
    rW, rS = axes['ranges']
    tau = axes['tau_points']
    w_out = axes['pade_freq'] if omega_out is None else np.asarray(omega_out, float)
    Ctw = grids.transform_weights(grids.COSINE_WT, tau, axes['freq_points'], *rW,
                                  warn=False)[0]
    C = grids.transform_weights(grids.COSINE_TW, tau, w_out, *rS, warn=False)[0]
    S = grids.transform_weights(grids.SINE_TW, tau, w_out, *rS, warn=False)[0]
    return Ctw, C, S


def diagonal_from_branches(sums, C, S):
    """Sigma^c_pp(i w) = -1/2 [(S^> + S^<)^T C^T + i (S^> - S^<)^T S^T],
    (nstates, nfreq) complex NumPy (imaginary_time.py:557-561)."""
    s = sums.cpu().numpy() if isinstance(sums, torch.Tensor) else np.asarray(sums)
    sl, sg = s[0], s[1]
    return -0.5 * ((sg + sl).T @ np.asarray(C).T + 1j * ((sg - sl).T @ np.asarray(S).T))


def self_energy_diagonal_space_time(X, V, eps, nocc, states, ntau=None,
                                    npade=grids.DEFAULT_NPADE,
                                    w0=grids.GL_FREQUENCY_SCALE, omega_out=None,
                                    timings=None, products='fp32'):
    """(sigma, axes): Sigma^c_pp(i w) (nstates, nfreq) complex of the states
    `states` on the Pade frequencies (or `omega_out`), from the collocation X
    (M, nocc + nvir), the grid Coulomb V (M, M) and the orbital energies.

    timings, a dict, receives the seconds of each stage (device synchronised)
    and 'sigma_kernel' ({'ms', 'launches', 'flops'} of the fused sums).
    products: the precision of the screening's products L^T chi0 L
    (`screened_correlation_frequencies`)."""
    dev = _device()
    eps = np.asarray(eps, float)
    states = np.atleast_1d(np.asarray(states, dtype=int))
    if states.min() < 0 or states.max() >= len(eps):
        raise ValueError('a state index is outside the orbitals')
    axes = grids.space_time_gw_grids(eps, nocc, ntau=ntau, npade=npade, w0=w0)
    clock = {}
    t0 = time.perf_counter()

    def lap(name):
        nonlocal t0
        if timings is not None:
            torch.cuda.synchronize()
            now = time.perf_counter()
            clock[name] = now - t0
            t0 = now
    Ctw, C, S = self_energy_transforms(axes, omega_out)
    lap('transforms')
    b64 = Chi0Builder(X, eps, nocc, mu=axes['mu'], precision='fp64', device=dev)
    Wc = screened_correlation_frequencies(b64, V, axes['grid'], products=products)
    del b64
    lap('screening')
    b = Chi0Builder(X, eps, nocc, mu=axes['mu'], precision='fp32', device=dev)
    Zt = screened_interaction_tau(Wc, Ctw)
    del Wc
    lap('to_tau')
    if timings is not None:
        b.profile(True)
    sums = b.sigma_sums(axes['tau_points'], Zt, states)
    lap('sigma')
    if timings is not None:
        clock['sigma_kernel'] = b.sigma_profile_read()
        b.profile(False)
        timings.update(clock)
    return diagonal_from_branches(sums, C, S), axes


def quasiparticle_energies_space_time(X, V, eps, nocc, states, xc_diagonal, ntau=None,
                                      npade=grids.DEFAULT_NPADE,
                                      w0=grids.GL_FREQUENCY_SCALE, with_z=False,
                                      timings=None):
    """G0W0 quasiparticle energies of `states` by the space-time route with the
    caller's <Sigma_x - v_xc>_pp (the serial core of
    `solve_qp_energy_space_time`, GW/space_time.py:631-849, with
    `finish_qp(xc_diagonal=...)`): Sigma^c(i w) on the Pade frequencies on the
    device, then the continuation and the largest-weight root (`qp`)."""
    from .qp import quasiparticle_energies
    sig, axes = self_energy_diagonal_space_time(X, V, eps, nocc, states, ntau=ntau,
                                                npade=npade, w0=w0, timings=timings)
    return quasiparticle_energies(sig, eps, nocc, states, axes['mu'], axes['pade_freq'],
                                  xc_diagonal, with_z=with_z)


def self_energy_diagonal_imaginary_time(X_mo, D, W_omega, eps, nocc, states, tau_points,
                                        omega_in, omega_out, mu=None, ranges=None,
                                        Wt_tau=None):
    """Sigma^c_pp(i omega_out), (nstates, nfreq) complex: the MO diagonal of
    `self_energy_matrix_imaginary_time` (GW/imaginary_time.py:403-533) for
    `states`, on the MO collocation X_mo (M, nmo) and an auxiliary factor D
    (M, naux) with W_omega = [I - chi0(i omega_in)]^-1 (nfreq, naux, naux), or
    the screened interaction already on the time axis, Wt_tau (ntau, naux,
    naux), whose W_omega is then ignored.  Zt = D Wt D^T is formed per time
    point (FP64) and the branch sums run in the fused FP32 kernel."""
    dev = _device()
    eps = np.asarray(eps, float)
    gap, span, mid = grids.particle_hole_window(eps, nocc)
    mu = mid if mu is None else float(mu)
    rW, rS = ranges or grids.sigma_fit_ranges(eps, nocc, mu=mu)
    tau_points = np.asarray(tau_points, float)
    Dd = _dev(D, dev)
    if Wt_tau is None:
        Ctw = grids.transform_weights(grids.COSINE_WT, tau_points, omega_in, *rW,
                                      warn=False)[0]
        Wm = _dev(W_omega, dev)
        eye = torch.eye(Wm.shape[-1], dtype=Wm.dtype, device=dev)
        Wt = torch.tensordot(torch.as_tensor(Ctw, device=dev), Wm - eye, dims=([1], [0]))
    else:
        Wt = _dev(Wt_tau, dev)
    M = Dd.shape[0]
    Zt = torch.empty((len(tau_points), M, M), dtype=torch.float32, device=dev)
    for k in range(len(tau_points)):
        Zt[k] = (Dd @ Wt[k] @ Dd.T).to(torch.float32)
    C = grids.transform_weights(grids.COSINE_TW, tau_points, omega_out, *rS, warn=False)[0]
    S = grids.transform_weights(grids.SINE_TW, tau_points, omega_out, *rS, warn=False)[0]
    b = Chi0Builder(X_mo, eps, nocc, mu=mu, precision='fp32', device=dev)
    sums = b.sigma_sums(tau_points, Zt, states)
    return diagonal_from_branches(sums, C, S)
