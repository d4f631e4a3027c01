"""A differentiable FP64 restatement of the qpGraph score in torch (test-side only).

Built from the oracle's graph tables (oracle/cpu_ref_qpgraph.cpp, cref_qpgraph_tables)
and following the reference's objective (src/core/qpadm/qpgraph_objective.hpp:24-190):
leaf weights from the path tables, centred weights, pair products P, cc = P^T G P,
ridge fudge tr(cc)/ne, and the edge lengths on a given active set A as
bl_A = C_AA^-1 q_A (the NNLS solution restricted to its passive set; the diagonal scaling
cancels).  torch's autograd of this function is the FP64 reference gradient the device
adjoint is checked against.
"""
import numpy as np
import torch


def leaf_weights(t, theta):
    """pwts (ne x npop) as a torch tensor for mixture weights theta (torch, float64)."""
    ne, npop = t["nedge_norm"], t["npop"]
    path_w = [torch.ones((), dtype=torch.float64) for _ in range(t["npath"])]
    for p, i in zip(t["pae_path"], t["pae_id"]):
        w = theta[(int(i) - 1) // 2]
        path_w[int(p)] = path_w[int(p)] * (w if int(i) % 2 == 1 else 1.0 - w)
    pw = [[torch.tensor(t["pwts0"][e, l], dtype=torch.float64) for l in range(npop)] for e in range(ne)]
    seen = set()
    n = len(t["pe_edge"])
    for a in range(n):
        e, l = int(t["pe_edge"][a]), int(t["pe_leaf"][a])
        if (e, l) in seen:
            continue
        seen.add((e, l))
        s = torch.zeros((), dtype=torch.float64)
        for b in range(a, n):
            if int(t["pe_edge"][b]) == e and int(t["pe_leaf"][b]) == l:
                s = s + path_w[int(t["pe_path"][b])]
        pw[e][l] = s
    return torch.stack([torch.stack(row) for row in pw])


def pair_products(t, theta):
    pw = leaf_weights(t, theta)
    base = t["base_leaf"]
    cols = [l for l in range(t["npop"]) if l != base]
    pc = pw[:, cols] - pw[:, base:base + 1]
    a = torch.as_tensor(np.asarray(t["cmb1"], np.int64))
    b = torch.as_tensor(np.asarray(t["cmb2"], np.int64))
    return (pc[:, a] * pc[:, b]).T  # (npair, ne)


def profiled_score(t, theta, f_obs, G, active, fudge=1e-4):
    """Score with bl_A = C_AA^-1 q_A on the active set (bool array over edges)."""
    P = pair_products(t, theta)
    G = torch.as_tensor(G, dtype=torch.float64)
    f = torch.as_tensor(f_obs, dtype=torch.float64)
    cc = P.T @ G @ P
    q = P.T @ (G @ f)
    ne = P.shape[1]
    C = cc + fudge * torch.trace(cc) / ne * torch.eye(ne, dtype=torch.float64)
    A = torch.as_tensor(np.nonzero(active)[0])
    bl = torch.zeros(ne, dtype=torch.float64)
    if len(A):
        bl = bl.index_put((A,), torch.linalg.solve(C[A][:, A], q[A]))
    res = f - P @ bl
    return res @ (G @ res), bl


def joint_score(t, theta, bl, f_obs, G):
    P = pair_products(t, theta)
    G = torch.as_tensor(G, dtype=torch.float64)
    res = torch.as_tensor(f_obs, dtype=torch.float64) - P @ bl
    return res @ (G @ res)
