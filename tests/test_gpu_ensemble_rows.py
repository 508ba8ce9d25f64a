"""Replicas of one engine with the widened rows (walls, angles, dihedrals,
exclusions, species fluxes): each replica of an ensemble evolves bit for bit
as an engine holding that replica alone, and per-replica term energies
match.  Cells, terms and fluxes are per replica in the device layout, so a
mix-up (a shared wall flag, a term list or flux accumulator crossing
replicas) breaks the equality."""
import numpy as np
import pytest

from mechanica_amd import workloads as W

pytestmark = pytest.mark.gpu


def replica_workload(i):
    w = W.molecules(n=5 + i, seed=40 + i)
    # walls in z; the molecules sit well inside the box
    w["boundaries"] = {"kinds": [2, 2, 2, 2, 4, 4]}
    w["species"] = np.random.default_rng(50 + i).random((len(w["x"]), 2)).astype(np.float32)
    w["fluxes"] = [(0, 0, 0, 0, 0, 0.5, 0.0, 0.0), (0, 0, 0, 1, 1, 0.2, 0.05, 0.0)]
    return w


def build_single(m, w):
    u = W.build_universe(m, w, seed=7)
    u.set_brick(2, 2, 2, 0, 0)
    return u


def test_ensemble_rows_match_singles(m):
    ws = [replica_workload(i) for i in range(3)]
    print("case: ensemble rows", [len(w["x"]) for w in ws], flush=True)
    singles = []
    for w in ws:
        u = build_single(m, w)
        u.step(6)
        s = u.state()
        singles.append((s, u.species(), u.term_energies(), u.observables()))
    w0 = ws[0]
    u = m.Universe(cutoff=w0["cutoff"], dt=w0["dt"], replicas=[(w["dim"], w["cells"]) for w in ws])
    for t in w0["types"]:
        u.add_type(t["mass"], t["radius"], t["dynamics"])
    keep = []
    for spec in w0["pots"]:
        p = W.make_potential(m, spec)
        u.bind(p, spec[2], spec[3])
        keep.append(p)
    u.set_seed(7)
    u.set_boundaries([2, 2, 2, 2, 4, 4])
    x = np.concatenate([w["x"] for w in ws])
    v = np.concatenate([w["v"] for w in ws])
    t = np.concatenate([w["type"] for w in ws])
    rep = np.concatenate([np.full(len(w["x"]), i, np.int32) for i, w in enumerate(ws)])
    u.load(x, v, t, replica=rep)
    base = np.cumsum([0] + [len(w["x"]) for w in ws])
    cat = lambda key, sub: np.concatenate([np.asarray(w[key][sub]) + base[i] for i, w in enumerate(ws)])
    bp = W.make_potential(m, w0["bonds"]["pot"])
    ap = W.make_potential(m, w0["angles"]["pot"])
    dp = W.make_potential(m, w0["dihedrals"]["pot"])
    keep += [bp, ap, dp]
    u.set_bonds(cat("bonds", "i"), cat("bonds", "j"), bp, 1e30)
    u.set_exclusions(cat("exclusions", "i"), cat("exclusions", "j"))
    u.set_angles(cat("angles", "i"), cat("angles", "j"), cat("angles", "k"), ap)
    u.set_dihedrals(cat("dihedrals", "i"), cat("dihedrals", "j"), cat("dihedrals", "k"), cat("dihedrals", "l"), dp)
    for fx in w0["fluxes"]:
        u.add_flux(*fx)
    u.set_species_values(np.concatenate([w["species"] for w in ws]))
    u.set_brick(2, 2, 2, 0, 0)
    u.step(6)
    s = u.state()
    sv, sq = u.species()
    cbase = 0
    for i, (si, (vi, qi), ti, oi) in enumerate(singles):
        a, b = base[i], base[i + 1]
        for k in ("x", "v", "f"):
            assert np.array_equal(s[k][a:b].view(np.uint32), si[k].view(np.uint32)), (i, k)
        assert np.array_equal(s["cell"][a:b] - cbase, si["cell"]), i
        assert np.array_equal(sv[a:b].view(np.uint32), vi.view(np.uint32)), i
        assert np.array_equal(sq[a:b].view(np.uint32), qi.view(np.uint32)), i
        te = u.term_energies(i)
        for k in te:
            assert abs(te[k] - ti[k]) <= 1e-12 * max(1.0, abs(ti[k])), (i, k)
        ob = u.observables(i)
        assert abs(ob["epot_bond"] - oi["epot_bond"]) <= 1e-12 * max(1.0, abs(oi["epot_bond"]))
        cbase += int(np.prod(ws[i]["cells"]))
