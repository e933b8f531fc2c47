"""High-level entry points: ``simspins`` and ``mesolve``.

Same signatures, defaults and plumbing as ``superspinsim/sugar.py``
(``simspins`` :385-584, ``mesolve`` :21-161); the simulation itself runs on
the HIP propagator through ``generate_simulator``.

Differences from the reference that do not change results:

* coefficient functions are plain Python callables (the reference
  numba-compiles them); ``mesolve`` builds its sampler as a closure instead
  of writing and importing a temporary module under ``~/.temp``;
* ``use_cuda`` is accepted for compatibility; the HIP device is always used.
"""

import numpy as np

from superspinsim_amd.generate_generators import (
    RotatingFrameError, _generate_lindbladian, _generate_superoperators,
    _generate_valid_indices, find_kernel, generate_atoms, make_rotating)


def _full(op):
    """numpy array of an operator; duck-types ``qutip.Qobj`` via ``full()``."""
    if isinstance(op, np.ndarray):
        return op
    if hasattr(op, "full") and callable(op.full):
        return np.asarray(op.full())
    return None


def prepare_simspins(coefficients, spins, spin_interactions,
                     group_interactions, number_of_exponentials=5,
                     number_of_fine_divisions=1,
                     number_of_quadratic_repeats=35, use_rotating=True,
                     use_residual=True, use_kernel=False, use_unitary=False,
                     use_hermitian=True, use_incoherent=True,
                     datatype=np.float64, verbose=False, use_cuda=True):
    """Everything ``simspins`` does before calling ``generate_simulator``.

    Returns ``(sampler, generators, vectorisation_map, kwargs)`` such that
    ``generate_simulator(sampler, generators, vectorisation_map, **kwargs)``
    is the simulator ``simspins`` runs (sugar.py:471-582).
    """
    operators, vectorisation_map = generate_atoms(
        spins, spin_interactions, group_interactions, use_unitary,
        use_hermitian, use_incoherent, verbose=verbose)
    if use_unitary and "unitary" not in operators:
        use_unitary = False

    unitary_kw = {}
    if not use_unitary:
        gens = operators["generators"]
    else:
        gens = operators["unitary"]
        unitary_kw = dict(operators["elimination"])
    first = "L0" if not use_unitary else "H0"
    generators = [gens[first], gens["Gx"], gens["Gy"], gens["Gz"], gens["Gr"]]

    kernel_kw = {}
    if use_kernel:
        full, image, projected = find_kernel(generators)
        if verbose:
            print(f"Old superoperator shape: {generators[0].shape}")
            print(f"New superoperator shape: {projected[0].shape}")
        generators = projected
        kernel_kw = {"full_projection": full, "image_projection": image}

    rotating_kw = {}
    if use_rotating:
        if np.sum(np.abs(generators[0])) == 0:
            use_rotating = False
        else:
            try:
                rotated, vr, ivr, doubles, singles = make_rotating(
                    generators[0], generators[1:])
                generators = rotated
                rotating_kw = {"vectors_real": vr, "inv_vectors_real": ivr,
                               "doubles": doubles, "singles": singles}
            except RotatingFrameError:
                print("Lindblad operator is not invertible.\n"
                      "Rotating frame cannot be used.")
                use_rotating = False

    sampler = _generate_lindbladian(coefficients, use_rotating=use_rotating,
                                    use_cuda=use_cuda)
    kwargs = dict(
        use_residual=use_residual,
        number_of_exponentials=number_of_exponentials,
        number_of_fine_divisions=number_of_fine_divisions,
        number_of_quartic_repeats=number_of_quadratic_repeats,
        use_rotating=use_rotating, **rotating_kw,
        use_kernel=use_kernel, **kernel_kw,
        use_unitary=use_unitary, **unitary_kw,
        use_hermitian=use_hermitian, use_cuda=use_cuda,
        datatype=datatype, verbose=verbose)
    return sampler, np.array(generators), vectorisation_map, kwargs


def simspins(coefficients, time_start, time_end, time_step, spins,
             spin_interactions, group_interactions, density_initial,
             number_of_exponentials=5, number_of_fine_divisions=1,
             number_of_quadratic_repeats=35, use_rotating=True,
             use_residual=True, use_kernel=False, use_unitary=False,
             use_cuda=True, use_hermitian=True, use_incoherent=True,
             datatype=np.float64, verbose=False):
    """Simulate a spin description (sugar.py:385-584).

    ``coefficients`` are four callables of time: x, y and z magnetic field
    (T) and the optical excitation as a fraction of saturation.  Returns
    ``(time (N,), density (N, N_h, N_h) complex128)``.
    """
    from superspinsim_amd.generate_simulator import generate_simulator
    sampler, generators, vmap, kwargs = prepare_simspins(
        coefficients, spins, spin_interactions, group_interactions,
        number_of_exponentials, number_of_fine_divisions,
        number_of_quadratic_repeats, use_rotating, use_residual, use_kernel,
        use_unitary, use_hermitian, use_incoherent, datatype, verbose,
        use_cuda)
    simulator = generate_simulator(sampler, generators, vmap, **kwargs)
    return simulator(density_initial, time_start, time_end, time_step)


# ---------------------------------------------------------------------------
# mesolve (sugar.py:21-382)
# ---------------------------------------------------------------------------

def _sort_operators(operators, label):
    message = (f"{label} must be numpy.ndarray, an [np.ndarray, function]"
               "list, or list of either.")
    quiescent, dependent, coefficients = [], [], []
    out = [quiescent, dependent, coefficients]
    if operators is None:
        return out
    single = _full(operators)
    if single is not None:
        quiescent.append(single)
        return out
    if type(operators) is not list:
        raise TypeError(message)
    if len(operators) < 1:
        return out
    if len(operators) == 2 and _pair(operators, dependent, coefficients):
        return out
    for op in operators:
        arr = _full(op)
        if arr is not None:
            quiescent.append(arr)
            continue
        if not _pair(op, dependent, coefficients):
            raise TypeError(message)
    return out


def _pair(pair, dependent, coefficients):
    if not isinstance(pair, (list, tuple)) or len(pair) != 2:
        return False
    arr = _full(pair[0])
    if arr is not None and callable(pair[1]):
        dependent.append(arr)
        coefficients.append(pair[1])
        return True
    return False


def _make_superoperators(ham_quiescent, ham_dependent, jump_quiescent,
                         jump_dependent, allowed, use_rotating,
                         use_hermitian=True):
    ops = {}
    for k, g in enumerate(ham_quiescent):
        ops[f"Hq{k}"] = g
    for k, g in enumerate(ham_dependent):
        ops[f"Ht{k}"] = g
    for k, g in enumerate(jump_quiescent):
        ops[f"Lq{k}"] = g
    for k, g in enumerate(jump_dependent):
        ops[f"Lt{k}"] = g
    ops = {k: np.asarray(v, dtype=complex) for k, v in ops.items()}
    if allowed is None:
        for group in (ham_quiescent, ham_dependent, jump_quiescent,
                      jump_dependent):
            if len(group) > 0:
                allowed = np.ones_like(np.real(group[0]))
                break
        else:
            raise TypeError("No operators specified.")
    valid_indices = _generate_valid_indices(allowed)
    supers = _generate_superoperators(ops, valid_indices, use_hermitian)

    def take(prefix):
        out, k = [], 0
        while f"{prefix}{k}" in supers:
            out.append(supers.pop(f"{prefix}{k}"))
            k += 1
        return out

    quiescent = take("Hq") + take("Lq")
    quiescent = sum(quiescent) if quiescent else None
    dependent = take("Ht") + take("Lt")
    if quiescent is None:
        if dependent:
            quiescent = np.zeros_like(dependent[0])
        else:
            raise TypeError("No operators defined.")

    rotating_kw = {}
    if use_rotating:
        try:
            dependent, vr, ivr, doubles, singles = make_rotating(quiescent,
                                                                 dependent)
            rotating_kw = {"vectors_real": vr, "inv_vectors_real": ivr,
                           "doubles": doubles, "singles": singles}
        except RotatingFrameError:
            print("Lindblad operator is not invertible.\n"
                  "Rotating frame cannot be used.")
            use_rotating = False
    if not use_rotating:
        dependent = [quiescent] + dependent
    return np.array(dependent), valid_indices, rotating_kw, use_rotating


def mesolve_coefficients(hamiltonian_coefficients, jump_coefficients,
                         use_rotating, intended_coefficients=False):
    """The sampler ``mesolve`` hands to ``generate_simulator``.

    The reference writes ``sample[k] = coefficient_k(t)`` for the Hamiltonian
    coefficients followed by ``sqrt`` of the jump coefficients
    (sugar.py:103-137, 244-257), whatever the frame.  In the rotating frame
    the generators are exactly the time-dependent ones, so slot k is
    coefficient k.  Outside it the generator list starts with the quiescent
    superoperator (:370-374): the reference then multiplies the quiescent
    part by the first coefficient function and shifts every other one by a
    slot (the last time-dependent generator gets coefficient 0).  That is
    reproduced by default; ``intended_coefficients=True`` gives the
    quiescent generator coefficient 1 and the functions their own slots.
    """
    funcs = list(hamiltonian_coefficients) + [
        (lambda f: (lambda t: np.sqrt(f(t))))(f) for f in jump_coefficients]
    offset = 1 if (intended_coefficients and not use_rotating) else 0

    def lindbladian(time, sample):
        if offset:
            sample[0] = 1.0
        for k, f in enumerate(funcs):
            sample[offset + k] = f(time)

    return lindbladian


def mesolve(H, rho0, ti, tf, dt, c_ops=None, allowed=None,
            use_rotating=True, use_hermitian=True, number_of_exponentials=5,
            number_of_fine_divisions=10, use_cuda=True,
            intended_coefficients=False):
    """QuTiP-like front end (sugar.py:21-161).

    Time-dependent jump operators ``[L, f]`` enter with coefficient
    ``sqrt(f(t))`` on the dissipator of L, as in the reference
    (sugar.py:123-137), i.e. the dissipator is scaled by sqrt(f), not f.
    Outside the rotating frame the coefficients fill the sampler's slots as
    the reference fills them (see ``mesolve_coefficients``; the keyword
    ``intended_coefficients`` selects the intended wiring instead).
    """
    from superspinsim_amd.generate_simulator import generate_simulator
    hq, hd, hc = _sort_operators(H, "H")
    hq = [sum(hq)] if hq else []
    lq, ld, lc = _sort_operators(c_ops, "c_ops")
    generators, vmap, rotating_kw, use_rotating = _make_superoperators(
        hq, hd, lq, ld, allowed, use_rotating, use_hermitian)
    lindbladian = mesolve_coefficients(hc, lc, use_rotating,
                                       intended_coefficients)

    simulator = generate_simulator(
        lindbladian, generators, vmap,
        number_of_exponentials=number_of_exponentials,
        number_of_fine_divisions=number_of_fine_divisions,
        use_rotating=use_rotating, **rotating_kw, use_cuda=use_cuda)
    return simulator(rho0, ti, tf, dt)
