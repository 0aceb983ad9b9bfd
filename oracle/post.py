"""CPU restatement of the docking post-processing (numpy).

TEST INFRASTRUCTURE ONLY: imported by tests/ to check the library's
post-processing (ud_cluster_modes / ud_postprocess), never by the product.

Follows Vina::global_search_gpu after do_docking (unidock/src/lib/vina.h:232-345)
on its no-refine path:
  add_to_output_container  coords.cpp:42-68 (find_closest coords.cpp:32-40,
                           rmsd_upper_bound coords.cpp:21-30)
  remove_redundant         vina.cpp:1465-1470
  score_gpu                vina.cpp:1245-1327 (num_tors_div,
                           conf_independent.cpp:153-156; w_rot internal,
                           vina.cpp:404)
  rmsd_lower/upper_bound   model.cpp:1021-1064

Pinned by tests/test_post.py: hand-built cases whose outcome is fixed by the
algorithm (a merge keeps the better of two close poses, a full container
replaces its worst, an energy of exactly 0 is skipped, translations give
RMSD = shift, permuted same-element atoms give a lower bound of 0) and the
closed form of num_tors_div.
"""
from __future__ import annotations

import numpy as np

EPS_FL = np.float32(1.1920928955078125e-07)


def rmsd_upper_bound(a, b):
    """sqrt(mean |a_i - b_i|^2) over equally ordered atoms, FP32 accumulation."""
    a = np.asarray(a, dtype=np.float32).reshape(-1, 3)
    b = np.asarray(b, dtype=np.float32).reshape(-1, 3)
    if len(a) == 0:
        return np.float32(0)
    acc = np.float32(0)
    for i in range(len(a)):
        d = a[i] - b[i]
        acc = np.float32(acc + np.float32(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]))
    return np.float32(np.sqrt(np.float32(acc / np.float32(len(a)))))


def rmsd_lower_bound(a, b, element):
    a = np.asarray(a, dtype=np.float64).reshape(-1, 3)
    b = np.asarray(b, dtype=np.float64).reshape(-1, 3)
    element = np.asarray(element)

    def one_way(x, y):
        if len(x) == 0:
            return 0.0
        acc = 0.0
        for i in range(len(x)):
            same = element == element[i]
            acc += np.min(np.sum((y[same] - x[i]) ** 2, axis=1))
        return np.sqrt(acc / len(x))

    return max(one_way(a, b), one_way(b, a))


def add_to_output_container(kept, cand, xyz, energy, min_rmsd, max_size):
    """One add_to_output_container step; `kept` is a list of candidate indices."""
    e = energy[cand]
    if -EPS_FL < e < EPS_FL:
        return kept
    closest, closest_r = -1, None
    for k, j in enumerate(kept):
        r = rmsd_upper_bound(xyz[cand], xyz[j])
        if k == 0 or r < closest_r:
            closest, closest_r = k, r
    kept = list(kept)
    if closest >= 0 and closest_r < min_rmsd:
        if e < energy[kept[closest]]:
            kept[closest] = cand
    elif len(kept) < max_size:
        kept.append(cand)
    elif kept and e < energy[kept[-1]]:
        kept[-1] = cand
    return sorted(kept, key=lambda j: energy[j])


def cluster(xyz, energy, min_rmsd, max_size):
    kept = []
    for t in range(len(energy)):
        kept = add_to_output_container(kept, t, xyz, energy, min_rmsd, max_size)
    return kept


def conf_independent(x, num_tors, rot_internal):
    weight = 0.1 * (rot_internal + 1.0)
    y = 1.0 + weight * num_tors / 5.0
    if abs(x) < EPS_FL:
        return 0.0
    return x / y


def postprocess(xyz, energy, inter, intra, num_tors, rot_internal, element, min_rmsd, max_modes):
    """One ligand: xyz[chain] heavy-atom coordinates, energy[chain] from the
    search, inter/intra callables giving the authentic-cap terms of a chain.
    Returns a list of dicts sorted by final energy."""
    first = cluster(xyz, energy, min_rmsd, max_modes)
    sub = cluster([xyz[j] for j in first], [energy[j] for j in first], min_rmsd, len(first))
    kept = [first[k] for k in sub]
    if not kept:
        return []
    intra_best = intra(kept[0])
    modes = []
    for j in kept:
        x = inter(j) + intra(j) - intra_best
        e = conf_independent(x, num_tors, rot_internal)
        modes.append(dict(chain=j, energy=e, inter=inter(j), intra=intra(j), search_energy=energy[j]))
    modes.sort(key=lambda m: m["energy"])
    best = xyz[modes[0]["chain"]]
    for m in modes:
        m["rmsd_ub"] = float(rmsd_upper_bound(xyz[m["chain"]], best))
        m["rmsd_lb"] = float(rmsd_lower_bound(xyz[m["chain"]], best, element))
    return modes
