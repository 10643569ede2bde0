"""Oracle (test infrastructure only): the reference's ExtCell tree step as a
dense solve.

Restates _dhs_bt._step (dendra/models/integrators/tree_bt.py:504-654) for
trees of one topology in node order: gates by the C oracle's hh_advance at
v = vi - ve0; membrane residual d = (J - G v) area + intra; residual
(d, -d - xg0 (ve0 - ve1), xg0 (ve0 - ve1) + xg1 (bath - ve1)); axial flux
f_i = e_i (vc_i - vc_p) (e_i vc_i at a root) subtracted at i and added at the
parent in ascending child order; block matrix D_i = main_i + ionic block,
A_ii = D_i + diag(e_i) + sum_c diag(e_c), A_ip = A_pi = -diag(e_i); the
increment from a dense numpy solve of the 3K system.  Python loops: small
trees.
"""

import numpy as np

import oracle


def run(ws, parent, hh_arr, dt, n_steps, state, *, lead=None, amp=None, wave=None,
        iamp=None, i_wave=None, record_every=0):
    main = np.asarray(ws["main_blocks"], dtype=np.float64)
    E = np.asarray(ws["g_to_parent"], dtype=np.float64)
    area = np.asarray(ws["area"], dtype=np.float64)
    xg = np.asarray(ws["xg"], dtype=np.float64)
    B, K = area.shape
    st = {k: np.array(v, dtype=np.float64) for k, v in state.items()}
    recs = []
    for s in range(n_steps):
        for b in range(B):
            V = st["vc"][b]
            A = np.zeros((3 * K, 3 * K))
            R = np.zeros(3 * K)
            for i in range(K):
                v = V[i, 0] - V[i, 1]
                m, h, n, G, J = oracle.hh_advance(v, dt, hh_arr, st["m"][b, i], st["h"][b, i],
                                                  st["n"][b, i])
                st["m"][b, i], st["h"][b, i], st["n"][b, i] = m, h, n
                g = G * area[b, i]
                d = (J - G * v) * area[b, i]
                if iamp is not None and i_wave is not None:
                    d += iamp[b, i] * i_wave[s]
                bath = lead[b, i] * amp[b] * wave[s] if lead is not None else 0.0
                shell = xg[b, i, 0] * (V[i, 1] - V[i, 2])
                R[3 * i:3 * i + 3] = [d, -d - shell, shell + xg[b, i, 1] * (bath - V[i, 2])]
                A[3 * i:3 * i + 3, 3 * i:3 * i + 3] += main[b, i] + np.array(
                    [[g, -g, 0.0], [-g, g, 0.0], [0.0, 0.0, 0.0]]) + np.diag(E[b, i])
            for i in range(K):
                p = parent[i]
                flux = E[b, i] * (V[i] - (V[p] if p >= 0 else 0.0))
                R[3 * i:3 * i + 3] -= flux
                if p >= 0:
                    R[3 * p:3 * p + 3] += flux
                    A[3 * p:3 * p + 3, 3 * p:3 * p + 3] += np.diag(E[b, i])
                    A[3 * i:3 * i + 3, 3 * p:3 * p + 3] -= np.diag(E[b, i])
                    A[3 * p:3 * p + 3, 3 * i:3 * i + 3] -= np.diag(E[b, i])
            st["vc"][b] = V + np.linalg.solve(A, R).reshape(K, 3)
        if record_every > 0 and (s + 1) % record_every == 0:
            recs.append(st["vc"][..., 0] - st["vc"][..., 1])
    return st, (np.stack(recs) if recs else None)
