"""Selected-CI driver: heat-bath selection, optionally guided by the learned selector.

The loop follows the reference's HCI usage (test_routines.py:442-470 and the
notebooks): solve, add_hci at a threshold, rebuild the operator, solve again.
With a `selector` (pyci_amd.mlp_selector), the network is retrained on the
current CI vector (target log10|c|) before it ranks the heat-bath pool taken at
eps/pool; the number appended per iteration equals what plain add_hci(eps)
would append, so the two runs are compared at equal determinant counts.
"""
import numpy as np

from pyci_amd._pyci import add_hci, sparse_op

__all__ = ["run_selection", "hartreefock_genci"]


def hartreefock_genci(nspatial, nalpha, nbeta):
    """Hartree-Fock determinant of a GenCI wave function over spin-orbitals
    alpha 0..n-1, beta n..2n-1 (the layout of utility.py:117-124)."""
    det = 0
    for i in range(nalpha):
        det |= 1 << i
    for i in range(nbeta):
        det |= 1 << (nspatial + i)
    return np.array([det], dtype=np.uint64)


def _solve(ham, wfn, c_prev, tol, precision):
    op = sparse_op(ham, wfn)
    c0 = np.zeros(len(wfn))
    if c_prev is None:
        c0[0] = 1.0
    else:
        c0[: c_prev.size] = c_prev
    es, cs = op.solve(n=1, c0=c0, tol=tol, precision=precision)
    return float(es[0]), cs[0]


def run_selection(ham, wfn, eps_schedule, selector=None, pool=5.0, score_mode=0, train_steps=300, batch=8192,
                  lr=1.0e-3, cmin=1.0e-10, seed=0, tol=1.0e-9, precision="f64", max_det=None, callback=None):
    """Runs one selection per threshold in `eps_schedule` (stopping early when
    nothing is added or once len(wfn) >= max_det). The first solve starts on
    determinant 0 and later ones on the previous vector padded with zeros, so
    the run stays in the symmetry sector of determinant 0. Returns a list of
    dicts (ndet, energy, added, eps, loss) after each selection; entry 0 is the
    starting space."""
    e, c = _solve(ham, wfn, None, tol, precision)
    traj = [dict(ndet=len(wfn), energy=e, added=0, eps=None, loss=None)]
    for it, eps in enumerate(eps_schedule):
        if max_det is not None and len(wfn) >= max_det:
            break
        loss = None
        if selector is not None:
            loss = float(selector.train(wfn, c, steps=train_steps, batch=batch, lr=lr, cmin=cmin, seed=seed + it))
        if selector is None:
            added = add_hci(ham, wfn, c, eps=eps)
        else:
            added = add_hci(ham, wfn, c, eps=eps, selector=selector, pool=pool, score_mode=score_mode)
        if added == 0:
            break
        e, c = _solve(ham, wfn, c, tol, precision)
        traj.append(dict(ndet=len(wfn), energy=e, added=int(added), eps=float(eps), loss=loss))
        if callback is not None:
            callback(traj[-1])
    return traj
