"""Drop-in ``generate_simulator`` running on the MI355X propagator.

Same signature, defaults and return values as
``superspinsim/generate_simulator.py:17-47`` / ``simulate`` (:1508-2202):

    simulate = generate_simulator(sampler, generators, vectorisation_map, ...)
    time, density = simulate(density_operator_initial, t0, t1, dt)

``time`` is (N_out,) float64 with ``N_out = int((t1 - t0) / dt)`` and
``time[i] = t0 + dt (i + 1)``; ``density`` is (N_out, N_h, N_h) complex128.

What runs where:

* host (numpy): flattening of rho0 through ``vectorisation_map``, the
  rotating-frame tables, tabulation of the user's ``sampler`` at every node
  time (``sampler(t, row)`` writes a zeroed row in place, as the reference's
  ``sample_kernel`` does), and the final basis changes / unflattening;
* device (``libssmx.so`` via ctypes): every propagator step of every output
  interval, in one kernel launch.

The reference builds each interval's operator by 2R residual squarings and a
serial prefix product; here the state is propagated with a Taylor action of
the same exponentials (degree chosen per call from a norm bound, truncation
error below 1e-16 per exponential), which equals the reference in exact
arithmetic.  ``use_residual=False`` runs the reference's own non-residual
arithmetic instead (``ssmx_propagate_dense``): every exponential is formed as
``(I + A / 4^R)`` squared 2R times, whose FP64 rounding drops the diagonal
(damping) terms exactly as the reference does (generate_simulator.py:534-541,
706-724).
"""

import ctypes

import numpy as np

from superspinsim_amd import _lib, quadratures
from superspinsim_amd.propagator import (
    Propagator, _ptr, device_tensor, eig_lane_factors, rotating_tables)

_f64 = np.float64


class UnsupportedSchemeError(NameError, ValueError):
    """``number_of_exponentials`` is not 1, 2, 3, 5 or 6.

    The reference raises ``NameError`` for these (no table is selected,
    generate_simulator.py:116-140); this subclass keeps that type.
    """


def _flatten(rho, vmap, use_hermitian, n_d):
    hilbert = rho.shape[0]
    k = np.arange(n_d)
    if use_hermitian:
        ys, xs, cs = vmap[:, 0], vmap[:, 1], vmap[:, 2]
    else:
        cs = k % 2
        xs = (k // 2) % hilbert
        ys = (k // 2) // hilbert
    vals = np.asarray(rho)[ys, xs]
    return np.where(cs == 0, np.real(vals), np.imag(vals)).astype(_f64)


def _unflatten(flat, vmap, use_hermitian, hilbert):
    n_t, n_d = flat.shape
    k = np.arange(n_d)
    if use_hermitian:
        ys, xs, cs = vmap[:, 0], vmap[:, 1], vmap[:, 2]
    else:
        cs = k % 2
        xs = (k // 2) % hilbert
        ys = (k // 2) // hilbert
    out = np.zeros((n_t, hilbert, hilbert, 2), dtype=_f64)
    out[:, ys, xs, cs] = flat
    if use_hermitian:
        off = ys != xs
        sign = np.where(cs[off] == 1, -1.0, 1.0)
        out[:, xs[off], ys[off], cs[off]] = flat[:, off] * sign
    return out[..., 0] + 1j * out[..., 1]


def tabulate(sampler, n_gen, n_out, fine, nodes, time_start, time_step):
    """Sampler rows at the reference node times: array (N_out, F, S, G).

    Node time of interval i, fine division r, node s is
    ``((t0 + (r dt) / F) + dt i) + (dt c_s) / F``
    (generate_simulator.py:161-167 with the call at :1956-1961).
    """
    S = nodes.size
    r = np.arange(fine, dtype=_f64)
    offset = time_start + (r * time_step) / fine            # (F,)
    i = np.arange(n_out, dtype=_f64)
    start = offset[None, :] + time_step * i[:, None]        # (N, F)
    ts = start[:, :, None] + (time_step * nodes) / fine     # (N, F, S)
    flat = ts.ravel()
    coef = np.zeros((flat.size, n_gen), dtype=_f64)
    for k in range(flat.size):
        sampler(flat[k], coef[k])
    return coef.reshape(n_out, fine, S, n_gen)


def plan_steps(flat, gen_norms, count=48):
    """Steps whose exponents plan the Taylor degree.

    ``flat`` (steps, S, K) holds the node coefficients of every step and
    ``gen_norms`` the 1-norms of the K generators they multiply ((K,), or
    (S, K) when each node has its own matrices).  Picks the ``count`` steps
    with the largest bound sum_{s,k} |c[s, k]| ||G_k||_1 on their
    exponent's norm, so that a short spike on a generator of large norm is
    planned for even where other coefficients are larger in magnitude, and
    ``count`` steps spread over the run.
    """
    bound = (np.abs(flat) * np.asarray(gen_norms, dtype=_f64)).sum(axis=(1, 2))
    count = min(count, flat.shape[0])
    return np.unique(np.concatenate([
        np.argsort(bound, kind="stable")[-count:],
        np.linspace(0, flat.shape[0] - 1, count).astype(np.int64)]))


def generate_simulator(
        sampler, generators, vectorisation_map=None, use_cuda=True,
        datatype=np.float64, use_residual=True, number_of_quartic_repeats=35,
        number_of_exponentials=2, number_of_fine_divisions=1,
        use_rotating=False, vectors_real=None, inv_vectors_real=None,
        doubles=None, singles=None, use_kernel=False, full_projection=None,
        image_projection=None, use_unitary=False, elimination=None,
        duplication=None, use_hermitian=True, verbose=False,
        use_cayley=False, tolerance=1e-16):
    """Build ``simulate(rho0, t0, t1, dt)`` for one system (see module doc).

    ``number_of_quartic_repeats`` is accepted for compatibility: the
    exponentials are evaluated to ``tolerance`` instead of by 2R
    squarings.  ``use_cuda`` is ignored (the HIP device is always used).
    """
    if use_cayley:
        raise TypeError("Cayley not implemented")
    if np.dtype(datatype) != np.float64:
        raise NotImplementedError("only float64 is supported")
    if number_of_exponentials not in quadratures.SCHEMES:
        raise UnsupportedSchemeError(
            f"number_of_exponentials={number_of_exponentials} has no CF "
            "table (valid: 1, 2, 3, 5, 6)")
    nodes, weights, label = quadratures.scheme(number_of_exponentials)
    if verbose:
        print(f"Using the {label} method with {nodes.size} nodes.")

    gens = np.asarray(generators).astype(_f64)
    n = gens.shape[1]
    if use_hermitian and vectorisation_map is None:
        use_hermitian = False
    if use_hermitian:
        vmap = np.asarray(vectorisation_map, dtype=np.int64)
        n_d = vmap.shape[0]
    else:
        vmap = None
        n_d = 2 * (n // 2) ** 2 if use_unitary else n
    F = int(number_of_fine_divisions)
    if F < 1:
        raise ValueError("number_of_fine_divisions must be >= 1")

    if use_rotating:
        vectors_real = np.array(vectors_real, dtype=_f64)
        inv_vectors_real = np.array(inv_vectors_real, dtype=_f64)
        doubles = np.array(doubles, dtype=_f64)
        singles = np.array(singles, dtype=_f64)
        if doubles.shape[0] == 0:
            doubles = doubles.reshape((0, 2))
    if use_kernel:
        image_projection = np.array(image_projection, dtype=_f64)
    if use_unitary:
        elimination = np.array(elimination, dtype=_f64)
        duplication = np.array(duplication, dtype=_f64)

    def simulate(density_operator_initial, time_start, time_end, time_step):
        torch = __import__("torch")
        N = int((time_end - time_start) / time_step)
        rho = np.asarray(density_operator_initial)
        hilbert = rho.shape[0]
        time = time_start + time_step * (np.arange(N) + 1)
        if N <= 0:
            return time, np.zeros((0, hilbert, hilbert), dtype=complex)
        h = time_step / F
        rho_flat = _flatten(rho, vmap, use_hermitian, n_d)

        coef = tabulate(sampler, gens.shape[0], N, F, nodes, time_start,
                        time_step)
        if not use_residual:
            states = _nonresidual_states(
                gens, weights, coef, rho_flat, h, number_of_quartic_repeats,
                use_rotating, doubles, singles, nodes, time_step, F,
                inv_vectors_real, use_kernel, image_projection, use_unitary)
            return time, _finish(states, n, rho_flat, vmap, use_hermitian,
                                 hilbert, use_rotating, vectors_real,
                                 inv_vectors_real, use_kernel,
                                 image_projection, use_unitary, elimination,
                                 duplication)
        # Generators whose coefficient never changes are summed into one
        # matrix with coefficient 1; all-zero coefficients are dropped.
        cols = coef.reshape(-1, coef.shape[-1])
        active = [g for g in range(cols.shape[1]) if np.any(cols[:, g] != 0)]
        const = {g: float(cols[0, g]) for g in active
                 if np.all(cols[:, g] == cols[0, g])}
        varying = [g for g in active if g not in const]
        if not active:
            const = {0: 0.0}
        if use_rotating:
            gen_all, d_step, s_step = rotating_tables(
                gens, doubles, singles, nodes, time_step, F, use_residual)
            eig = eig_lane_factors(d_step, s_step, n, use_residual)
            parts = []
            if const:
                parts.append(sum(v * gen_all[:, g]
                                 for g, v in const.items())[:, None])
            if varying:
                parts.append(gen_all[:, varying])
            mats = np.concatenate(parts, axis=1)
        else:
            parts = []
            if const:
                parts.append(sum(v * gens[g] for g, v in const.items())[None])
            if varying:
                parts.append(gens[varying])
            mats = np.concatenate(parts, axis=0)
            eig = None
        ones = [np.ones(coef.shape[:-1] + (1,))] if const else []
        coef = np.ascontiguousarray(np.concatenate(
            ones + [coef[..., varying]], axis=-1))
        prop = Propagator(mats, weights, nodes, rotating=use_rotating,
                          eig=eig)
        flat = coef.reshape(-1, coef.shape[2], coef.shape[3])  # (N*F, S, K)
        pick = plan_steps(flat, prop.gen_norms)
        m, s = prop.plan(flat[pick], h, tol=tolerance)
        if verbose:
            print(f"Taylor degree {m}, {s} substeps, {prop.ell.slots} ELL "
                  f"slots, nnz {prop.ell.nnz}")

        # unitary mode propagates every column of the identity: the
        # unitary's real interleaved form at each output (operator mode)
        x0 = _initial_rows(rho_flat, n, use_rotating, inv_vectors_real,
                           use_kernel, image_projection, use_unitary)
        n_inst = x0.shape[0]
        d_rho = device_tensor(x0)
        d_coef = device_tensor(coef)
        out = prop.run(n_inst=n_inst, fine=F, n_out=N, h=h, t0=time_start,
                       dt=time_step, rho0=d_rho, coef=d_coef,
                       coef_inst_stride=0, taylor=m, substeps=s,
                       read="full")
        torch.cuda.synchronize()
        states = out.cpu().numpy()                           # (I, N, n)
        return time, _finish(states, n, rho_flat, vmap, use_hermitian,
                             hilbert, use_rotating, vectors_real,
                             inv_vectors_real, use_kernel, image_projection,
                             use_unitary, elimination, duplication)

    return simulate


def _initial_rows(rho_flat, n, use_rotating, inv_vectors_real, use_kernel,
                  image_projection, use_unitary):
    """States the device propagates: the identity's columns (unitary mode,
    operator form) or rho0 projected onto the image (kernel mode) and moved
    into the eigenbasis (rotating frame)."""
    if use_unitary:
        x0 = np.eye(n, dtype=_f64)
    else:
        x0 = rho_flat[None, :] if not use_kernel else \
            (image_projection.T @ rho_flat)[None, :]
    if use_rotating and not use_unitary:
        x0 = x0 @ inv_vectors_real.T
    return np.ascontiguousarray(x0)


def _finish(states, n, rho_flat, vmap, use_hermitian, hilbert, use_rotating,
            vectors_real, inv_vectors_real, use_kernel, image_projection,
            use_unitary, elimination, duplication):
    """Device states (I, N, n) to densities: the global sandwiches of the
    reference (rotating V, kernel P, unitary lift; :2030-2105) and the
    unflattening (:2156-2191)."""
    if use_unitary:
        return _unitary_density(states, n, rho_flat, elimination,
                                duplication, vmap, use_hermitian, hilbert,
                                use_rotating, vectors_real, inv_vectors_real)
    x = states[0]                                            # (N, n)
    if use_rotating:
        x = x @ vectors_real.T
    if use_kernel:
        base = rho_flat - image_projection @ (image_projection.T @ rho_flat)
        x = x @ image_projection.T + base[None, :]
    return _unflatten(x, vmap, use_hermitian, hilbert)


def _nonresidual_states(gens, weights, coef, rho_flat, h, repeats,
                        use_rotating, doubles, singles, nodes, time_step, F,
                        inv_vectors_real, use_kernel, image_projection,
                        use_unitary):
    """States (I, N, n) of the reference's non-residual arithmetic on the
    device: every generator and coefficient as tabulated (no folding, so
    the assembly sums the same terms in the same order as the reference)."""
    torch = __import__("torch")
    n = gens.shape[1]
    if n > _lib.MAX_N:
        raise NotImplementedError(
            f"use_residual=False is supported for n <= {_lib.MAX_N}")
    x0 = _initial_rows(rho_flat, n, use_rotating, inv_vectors_real,
                       use_kernel, image_projection, use_unitary)
    N, _, S, G = coef.shape
    flags, nd, ns, eig = 0, 0, 0, None
    mats = gens
    if use_rotating:
        mats, d_step, s_step = rotating_tables(gens, doubles, singles, nodes,
                                               time_step, F, residual=False)
        flags = _lib.SSMX_ROTATING
        nd, ns = d_step.shape[0], s_step.size
        eig = device_tensor(np.concatenate([d_step.ravel(), s_step]))
    d_gens = device_tensor(np.ascontiguousarray(mats, dtype=_f64))
    d_w = device_tensor(np.ascontiguousarray(weights, dtype=_f64))
    d_coef = device_tensor(np.ascontiguousarray(coef, dtype=_f64))
    d_rho = device_tensor(x0)
    n_inst = x0.shape[0]
    out = torch.empty((n_inst, N, n), dtype=torch.float64, device="cuda")
    desc = _lib.DenseDesc(
        n=n, n_inst=n_inst, n_gen=G, n_exp=weights.shape[0], n_node=S,
        fine=F, n_out=N, squarings=2 * int(repeats), flags=flags,
        n_doubles=nd, n_singles=ns, reserved=0, coef_inst_stride=0, h=h,
        scale=4.0 ** int(repeats))
    stream = torch.cuda.current_stream().cuda_stream
    _lib.check(_lib.load().ssmx_propagate_dense(
        ctypes.byref(desc), _ptr(d_gens), _ptr(d_w), _ptr(d_coef),
        _ptr(d_rho), _ptr(eig) if eig is not None else None, _ptr(out),
        ctypes.c_void_p(stream)))
    torch.cuda.synchronize()
    return out.cpu().numpy()


def _unitary_density(columns, n, rho_flat, elimination, duplication, vmap,
                     use_hermitian, hilbert, use_rotating, vectors_real,
                     inv_vectors_real):
    """Lift propagated unitary columns to rho(t) (:1394-1504, 2074-2116).

    ``columns[k, t]`` is the propagator at output t applied to unit vector
    k, i.e. column k of the real interleaved unitary form.  The lift follows
    the reference: element (y, x) of the superoperator, with
    y = y_out * N + y_in, is conj(U[y_out, x_out]) U[y_in, x_in] in the real
    2x2 form [[r, i], [-i, r]], then rho_t = E (lift) D rho0.
    """
    te = np.transpose(columns, (1, 2, 0))                    # (N_out, n, n)
    if use_rotating:
        te = vectors_real @ te @ inv_vectors_real
    half = n // 2
    ur = te[:, 0::2, 0::2]
    ui = te[:, 1::2, 0::2]
    u = ur + 1j * ui                                         # (N_out, N, N)
    lifted = np.einsum("tac,tbd->tabcd", np.conj(u), u).reshape(
        te.shape[0], half * half, half * half)
    lr, li = lifted.real, lifted.imag
    size = half * half
    full = np.zeros((te.shape[0], 2 * size, 2 * size), dtype=_f64)
    full[:, 0::2, 0::2] = lr
    full[:, 1::2, 1::2] = lr
    full[:, 1::2, 0::2] = -li
    full[:, 0::2, 1::2] = li
    flat = np.einsum("ij,tjk,kl,l->ti", elimination, full, duplication,
                     rho_flat)
    return _unflatten(flat, vmap, use_hermitian, hilbert)
