"""Stage-by-stage pins of the potential-table builder.

The builder (mechanica_amd/csrc/mx_potential.cpp) restates potential_init
and its helpers (src/mdcore/src/MxPotential.cpp:1690-1932 potential_init,
:1949-2029 potential_getfp, :2032-2112 potential_getfp_fixend, :2139-2260
potential_getcoeffs, :2277-2340 potential_getalpha).  The reference cannot be
built here, so each stage is checked against an independent numpy
restatement of the mathematics it computes, on the potentials the configured
workloads use.  tests/native/table_probe.cpp compiles the builder's own
source into a small host library that exposes those stages; the first test
shows that it reproduces libmxg.so's tables bit for bit.

  * closed forms: f, f' and f^(6) against the analytic derivatives (the
    sixth by falling factorials, not by the builder's expressions) and
    against central differences;
  * node exponent: the minimiser of max_j |f6(t_j)| / m'(t_j)^6 over
    [0, 2] (m(t) = t (e + (1 - e) t) is the interval map in t = (r-a)/(b-a))
    found by scipy's bounded Brent search, to 1e-4;
  * interval map: alpha encodes m; the edges sit where n m(t) = i;
  * the seven cardinal constants: the discrete weighted inner products of
    the cubic Hermite cardinal functions on the N - 2 interior
    Chebyshev-Lobatto nodes;
  * edge slopes: the least-squares problem they solve, assembled directly
    (dense, numpy lstsq, no cardinal constants, no Thomas sweep);
  * interval fits: the Hermite cubic's Chebyshev coefficients (numpy), the
    degree-4/5 least-squares corrections against phi, and the monomial
    conversion (numpy cheb2poly);
  * the builder's error estimate recomputed from the stored coefficients,
    the tolerance met at its nodes (<= tol), n minimal (n - 1 misses tol);
  * the continuation below a: the quadratic Taylor expansion of interval 1;
  * the float/double choice of fabs/sqrt in the domain stretch and the map
    correction does not change a single bit for the configured potentials.

The reference's phi (potential_getcoeffs, MxPotential.cpp:2210-2214) puts the
interval's left value where the Chebyshev-Lobatto sum has the right one for
odd j.  That shifts phi_1, phi_3 and phi_5 by the same amount, and the odd
correction's direction (-2, 3, -1) is orthogonal to a uniform shift, so the
tables are the same either way (checked below); the restatement here follows
the reference's order.
"""
import ctypes as C
import math
import os
import subprocess

import numpy as np
import pytest
from numpy.polynomial import chebyshev as npcheb
from scipy.optimize import minimize_scalar

from mechanica_amd import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "mechanica_amd", "csrc")
PROBE_SRC = os.path.join(ROOT, "tests", "native", "table_probe.cpp")
FEPS = float(np.finfo(np.float32).eps)
NS = 100  # samples per interval (potential_N)

# (kind, closed-form parameters, factory arguments, tol) of the configured
# potentials: argon LJ, lattice Morse, cell-sort / adjoint Morse (d = 3, 1 on
# [1e-4, 2.5]), lattice bonds, molecule angle and dihedral.
LJ = (0, (W.LJ_A, W.LJ_B), (0.275, 1.0, W.LJ_A, W.LJ_B), 1e-3)
MORSE = (1, (1.0, 6.0), (1.0, 6.0, 1.0, 1e-4, 2.0), 1e-3)
MORSE3 = (1, (3.0, 6.0), (3.0, 6.0, 1.0, 1e-4, 2.5), 1e-3)
MORSE1 = (1, (1.0, 6.0), (1.0, 6.0, 1.0, 1e-4, 2.5), 1e-3)
HARM = (2, (100.0, 1.0), (0.0, 2.0, 100.0, 1.0), 0.02)
TH0 = math.radians(110.0)
MOL_A, MOL_B = 4 * 0.5 * 0.25 ** 12, 4 * 0.5 * 0.25 ** 6
LJMOL = (0, (MOL_A, MOL_B), (0.1, 1.0, MOL_A, MOL_B), 1e-4)
BOND = (2, (100.0, 0.3), (0.0, 0.6, 100.0, 0.3), 1e-4)
ANGLE = (3, (20.0, TH0), (TH0 - math.radians(60), TH0 + math.radians(60), 20.0, TH0), 1e-4)
DIHED = (4, (1.0, 3, 0.0), (1.0, 3, 0.0), 1e-4)
CONFIGS = {"lj": LJ, "morse": MORSE, "morse3": MORSE3, "morse1": MORSE1, "harmonic": HARM, "ljmol": LJMOL,
           "bond": BOND, "angle": ANGLE, "dihedral": DIHED}
WITH_F6 = ("lj", "morse", "morse3", "morse1", "ljmol")


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _f(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _i(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


@pytest.fixture(scope="module")
def probe(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("probe") / "table_probe.so")
    # The host compiler hipcc uses for libmxg.so (clang).  g++ -O3 rounds one
    # map coefficient differently (one ulp of alpha[0] for the argon table);
    # clang at any level and g++ up to -O2 agree with the library.
    clang = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "llvm", "bin", "clang++")
    cxx, opt = (clang, "-O3") if os.path.exists(clang) else ("g++", "-O2")
    cmd = [cxx, opt, "-std=c++17", "-ffp-contract=off", "-fno-fast-math", "-fPIC", "-shared", "-I", CSRC,
           PROBE_SRC, "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    lib = C.CDLL(out)
    lib.probe_node_exponent.restype = C.c_double
    lib.probe_node_exponent.argtypes = [C.c_int, C.POINTER(C.c_double), C.c_double, C.c_double]
    lib.probe_edge_slopes.argtypes = [C.c_int, C.POINTER(C.c_double), C.c_double, C.c_double, C.c_int,
                                      C.POINTER(C.c_float), C.POINTER(C.c_double), C.c_int]
    lib.probe_fix_first.argtypes = [C.POINTER(C.c_float), C.c_float]
    lib.probe_place_edges.argtypes = [C.POINTER(C.c_float), C.c_float, C.c_float, C.c_int, C.POINTER(C.c_float)]
    lib.probe_build.argtypes = [C.c_int, C.POINTER(C.c_double), C.c_double, C.POINTER(C.c_int),
                                C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_int, C.POINTER(C.c_float),
                                C.POINTER(C.c_double), C.POINTER(C.c_float), C.POINTER(C.c_int),
                                C.POINTER(C.c_float), C.c_int, C.POINTER(C.c_int)]
    assert lib.probe_sample_count() == NS
    return lib


class Probe:
    """The builder's stages for one configured potential."""

    def __init__(self, lib, name):
        self.lib = lib
        self.name = name
        self.kind, par, args, self.tol = CONFIGS[name]
        self.par = np.array(par, dtype=np.float64)
        self.args = np.array(args, dtype=np.float64)

    def eval(self, which, r):
        r = np.ascontiguousarray(r, dtype=np.float64)
        out = np.zeros_like(r)
        rc = self.lib.probe_eval(self.kind, _d(self.par), which, len(r), _d(r), _d(out))
        return None if rc else out

    def f(self, r):
        return self.eval(0, r)

    def fp(self, r):
        return self.eval(1, r)

    def build(self):
        n = C.c_int()
        alpha = np.zeros(4, np.float32)
        coef = np.zeros(8 * 4096, np.float32)
        ab = np.zeros(2, np.float32)
        ex = C.c_double()
        al = np.zeros(3, np.float32)
        tn = np.zeros(64, np.int32)
        te = np.zeros(64, np.float32)
        nt = C.c_int()
        rc = self.lib.probe_build(self.kind, _d(self.args), self.tol, C.byref(n), _f(alpha), _f(coef), len(coef),
                                  _f(ab), C.byref(ex), _f(al), _i(tn), _f(te), 64, C.byref(nt))
        assert rc == 0, rc
        return {"n": n.value, "alpha": alpha, "c": coef[:(n.value + 1) * 8].copy(), "a": ab[0], "b": ab[1],
                "ex": ex.value, "al": al, "tried": list(zip(tn[:nt.value].tolist(), te[:nt.value].tolist()))}

    def edges(self, al, a, b, n):
        x = np.zeros(n + 1, np.float32)
        assert self.lib.probe_place_edges(_f(np.asarray(al, np.float32)), float(a), float(b), n, _f(x)) == 0
        return x

    def free_ends(self):
        return self.fp(np.array([0.5])) is None

    def slopes(self, x):
        n = len(x) - 1
        s = np.zeros(n + 1)
        free = self.free_ends()
        fpa = 0.0 if free else float(self.fp(np.array([float(x[0])]))[0])
        fpb = 0.0 if free else float(self.fp(np.array([float(x[n])]))[0])
        self.lib.probe_edge_slopes(self.kind, _d(self.par), fpa, fpb, n, _f(x), _d(s), int(free))
        return s

    def fit(self, x):
        n = len(x) - 1
        out = np.zeros(n * 8, np.float32)
        err = C.c_float()
        self.lib.probe_fit(self.kind, _d(self.par), _f(x), n, _f(out), C.byref(err))
        return out.reshape(n, 8), err.value


def horner_monomial(c, s):
    """The chunk's polynomial at local s (c[2] highest power ... c[7])."""
    v = np.zeros_like(s, dtype=np.float64) + float(c[2])
    for k in range(3, 8):
        v = v * s + float(c[k])
    return v


def builder_error(P, c, x):
    """The builder's error estimate recomputed from stored coefficients: at
    the nodes cos(k pi / N), k = 1..N-1 (float-rounded, as its table of
    cosines), |f - p| relative to max(max |f| there, 1), worst interval."""
    k = np.arange(1, NS)
    s32 = np.cos(k * np.pi / NS).astype(np.float32).astype(np.float64)
    worst = 0.0
    for i in range(len(x) - 1):
        mid = np.float64(np.float32(x[i] + x[i + 1]) / np.float32(2))
        half = np.float64(np.float32(x[i + 1] - x[i]) / np.float32(2))
        fs = P.f(mid + half * np.cos(k * np.pi / NS))
        e = np.abs(fs - horner_monomial(c[i], s32)).max() / max(np.abs(fs).max(), 1.0)
        worst = max(worst, e)
    return worst


# ---------------------------------------------------------------- the probe


def test_probe_reproduces_library_tables(probe, m):
    """The probe compiles the same builder: n, alpha and every coefficient
    equal libmxg.so's (which the GPU and the oracle consume)."""
    lib_pots = {
        "lj": lambda: m.Potential.lennard_jones_12_6(0.275, 1.0, W.LJ_A, W.LJ_B, 1e-3),
        "morse": lambda: m.Potential.morse(d=1.0, a=6.0, r0=1.0, min=1e-4, max=2.0, tol=1e-3),
        "morse3": lambda: m.Potential.morse(d=3.0, a=6.0, r0=1.0, min=1e-4, max=2.5, tol=1e-3),
        "harmonic": lambda: m.Potential.harmonic(100.0, 1.0, min=0.0, max=2.0, tol=0.02),
        "ljmol": lambda: m.Potential.lennard_jones_12_6(0.1, 1.0, MOL_A, MOL_B, 1e-4),
        "bond": lambda: m.Potential.harmonic(100.0, 0.3, min=0.0, max=0.6, tol=1e-4),
        "angle": lambda: m.Potential.harmonic_angle(20.0, TH0, min=TH0 - math.radians(60), max=TH0 + math.radians(60),
                                                    tol=1e-4),
        "dihedral": lambda: m.Potential.harmonic_dihedral(1.0, 3, 0.0, tol=1e-4),
    }
    for name, mk in lib_pots.items():
        p = mk()
        b = Probe(probe, name).build()
        assert p.n == b["n"], name
        assert np.array_equal(p.alpha.view(np.uint32), b["alpha"].view(np.uint32)), name
        assert np.array_equal(p.coefficients.view(np.uint32), b["c"].view(np.uint32)), name


# ---------------------------------------------------------------- closed forms


@pytest.mark.parametrize("name", ["lj", "morse3"])
def test_closed_forms(probe, name):
    P = Probe(probe, name)
    r = np.linspace(0.3, 2.4, 400) if name != "lj" else np.linspace(0.28, 1.0, 400)
    if name == "lj":
        A, B = P.par
        e = A / r ** 12 - B / r ** 6
        d1 = -12 * A / r ** 13 + 6 * B / r ** 7
        d6 = A * math.prod(range(12, 18)) / r ** 18 - B * math.prod(range(6, 12)) / r ** 12
    else:
        d, a = P.par
        u = np.exp(-a * (r - 1))
        e = d * (u * u - 2 * u)
        d1 = d * (-2 * a * u * u + 2 * a * u)
        d6 = d * (64 * a ** 6 * u * u - 2 * a ** 6 * u)
    assert np.allclose(P.f(r), e, rtol=1e-13, atol=1e-300)
    assert np.allclose(P.fp(r), d1, rtol=1e-12, atol=1e-12 * np.abs(d1).max())
    assert np.allclose(P.eval(2, r), d6, rtol=1e-12, atol=1e-12 * np.abs(d6).max())
    h = 1e-6
    cd = (P.f(r + h) - P.f(r - h)) / (2 * h)
    assert np.allclose(P.fp(r), cd, rtol=1e-6, atol=1e-6 * np.abs(d1).max())


# ---------------------------------------------------------------- node exponent


def stretched(a, b):
    """potential_init's stretch of [a, b] (MxPotential.cpp:1717-1719) in
    float, and with the double overloads of fabs and sqrt."""
    a32, b32 = np.float32(a), np.float32(b)
    s = np.float32(np.sqrt(np.float32(FEPS)))
    fl = (np.float32(a32 - np.float32(abs(a32)) * s), np.float32(b32 + np.float32(abs(b32)) * s))
    dbl = (np.float32(float(a32) - abs(float(a32)) * math.sqrt(FEPS)),
           np.float32(float(b32) + abs(float(b32)) * math.sqrt(FEPS)))
    return fl, dbl


@pytest.mark.parametrize("name", WITH_F6)
def test_node_exponent_is_the_bound_minimiser(probe, name):
    """potential_getalpha: the exponent e of the node distribution minimises
    max_j |f6(a + (b - a) t_j)| / (e + 2 (1 - e) t_j)^6 over [0, 2] on the
    samples t_j = (j + 1) / (N + 1); that bound is quasi-convex in e (each
    term is monotone), so the minimiser is unique."""
    P = Probe(probe, name)
    B = P.build()
    a, b = float(B["a"]), float(B["b"])
    t = (np.arange(NS) + 1.0) / (NS + 1)
    g = np.abs(P.eval(2, a + (b - a) * t))

    def bound(e):
        return (g / (e + 2 * (1 - e) * t) ** 6).max()

    res = minimize_scalar(bound, bounds=(0.0, 2.0), method="bounded", options={"xatol": 1e-10})
    ex = B["ex"]
    assert abs(ex - res.x) <= 1e-4, (ex, res.x)
    assert ex == probe.probe_node_exponent(P.kind, _d(P.par), a, b)
    # The stored map carries the same exponent: al2 = (1 - e) / (a - b)^2.
    w = 1.0 / (a - b) ** 2
    assert abs((1.0 - float(B["al"][2]) / w) - ex) <= 1e-5 * max(1.0, abs(ex))


def test_node_exponent_without_sixth_derivative(probe):
    """A polynomial of degree below six (f6 = 0) takes the uniform map."""
    assert probe.probe_node_exponent(2, _d(np.array([100.0, 1.0])), 0.0, 2.0) == 1.0


# ---------------------------------------------------------------- interval map


@pytest.mark.parametrize("name", sorted(CONFIGS))
def test_interval_map_and_edges(probe, name):
    """alpha is m(t) = t (e + (1 - e) t) in t = (r - a) / (b - a), shifted
    right by 2 eps (|al0| + |al1| + |al2|) (r - a) / (b - a); the edges of the
    chosen table sit where n m = i."""
    P = Probe(probe, name)
    B = P.build()
    a, b, ex, n = float(B["a"]), float(B["b"]), B["ex"], B["n"]
    al = B["al"].astype(np.float64)
    corr = 2 * FEPS * np.abs(B["al"]).astype(np.float64).sum()
    r = np.linspace(a, b, 257)
    t = (r - a) / (b - a)
    m_expected = t * (ex + (1 - ex) * t) - corr * t
    m_stored = al[0] + r * (al[1] + r * al[2])
    scale = np.abs(al).sum() * max(abs(a), abs(b), 1.0) ** 2
    assert np.abs(m_stored - m_expected).max() <= 8 * FEPS * scale
    # Scaled by n with the continuation interval in front.
    assert np.allclose(B["alpha"][:3].astype(np.float64), np.array([al[0] * n + 1, al[1] * n, al[2] * n]),
                       rtol=2 * FEPS, atol=2 * FEPS * scale * n)
    # Edges from the stored chunks (centre, 1/half-width): n m(x_i) = i to
    # within the Newton stop (10 eps n) plus the float rounding of x.
    c = B["c"].reshape(n + 1, 8)[1:]
    half = 1.0 / c[:, 1].astype(np.float64)
    lo, hi = c[:, 0] - half, c[:, 0] + half
    x = np.concatenate([lo[:1], hi])
    dm = np.abs(al[1] + 2 * x * al[2]) * n
    tol_i = 10 * FEPS * n + dm * (np.abs(x) + 1e-30) * 4 * FEPS + 8 * FEPS * scale * n
    assert np.all(np.abs(n * (al[0] + x * (al[1] + x * al[2])) - np.arange(n + 1)) <= tol_i)
    assert np.all(np.diff(x) > 0)
    # Neighbouring chunks share their edge.
    assert np.allclose(hi[:-1], lo[1:], rtol=4 * FEPS, atol=4 * FEPS * max(abs(a), abs(b)))


@pytest.mark.parametrize("name", sorted(CONFIGS))
def test_overload_choice_is_immaterial(probe, name):
    """The reference's unqualified fabs/sqrt (MxPotential.cpp:1718-1719,
    1736) resolve through headers absent here.  The builder takes the float
    overloads (std::fabs/std::sqrt of float); the double ones give the same
    stretched domain and the same map bits for every configured potential."""
    P = Probe(probe, name)
    B = P.build()
    if name == "angle":
        a0, b0 = math.cos(P.args[1]), math.cos(P.args[0])
    elif name == "dihedral":
        a0, b0 = -1.0, 1.0
    elif name.startswith("morse"):
        a0, b0 = P.args[3], P.args[4]
    else:
        a0, b0 = P.args[0], P.args[1]
    fl, dbl = stretched(a0, b0)
    assert (fl[0], fl[1]) == (B["a"], B["b"])
    assert (dbl[0], dbl[1]) == (B["a"], B["b"])
    al = B["al"]
    # The map correction w = 2 eps (|al0| + |al1| + |al2|) in float or in
    # double, applied as the builder does (double arithmetic, float store).
    # The unscaled map as potential_init forms it (MxPotential.cpp:1728-1732):
    # a, b floats, a - b, a * a and a + b rounded to float, the rest double.
    a32, b32, ex = B["a"], B["b"], B["ex"]
    a, b = float(a32), float(b32)
    amb = float(np.float32(a32 - b32))
    w0 = (1.0 / amb) ** 2
    raw = np.array([np.float32(float(np.float32(a32 * a32)) * w0 - ex * b * a * w0),
                    np.float32(-2 * a * w0 + ex * float(np.float32(a32 + b32)) * w0),
                    np.float32(w0 - ex * w0)], np.float32)
    w_f = float(np.float32(2 * np.float32(FEPS)) * np.float32(np.abs(raw).sum(dtype=np.float32)))
    w_d = 2 * FEPS * float(np.abs(raw.astype(np.float64)).sum())
    for w in (w_f, w_d):
        got = np.array([np.float32(float(raw[0]) - w * a / amb), np.float32(float(raw[1]) + w / amb), raw[2]],
                       np.float32)
        assert np.array_equal(got.view(np.uint32), al.view(np.uint32)), (name, w)


# ---------------------------------------------------------------- slopes


def hermite_cardinals():
    return {
        "wl0": lambda s: (1 - s) ** 2 * (2 + s) / 4,
        "wr0": lambda s: (1 + s) ** 2 * (2 - s) / 4,
        "wl1": lambda s: (1 - s) ** 2 * (1 + s) / 4,
        "wr1": lambda s: -(1 + s) ** 2 * (1 - s) / 4,
    }


def lobatto_interior():
    k = np.arange(1, NS - 1)
    s = np.cos(k * np.pi / (NS - 1))
    return s, 1.0 / np.sqrt(1.0 - s * s)


def test_cardinal_constants(probe):
    """The seven constants of the slope system are the discrete inner
    products <u, v> = sum_k w_k u(s_k) v(s_k) / (N - 2) of the cubic Hermite
    cardinal functions over the interior Chebyshev-Lobatto nodes
    s_k = cos(k pi / (N - 1)), w_k = 1 / sqrt(1 - s_k^2) -- the same
    quadrature as the slope projections, so the system is the exact normal
    equations of that discrete least-squares problem."""
    out = np.zeros(7)
    probe.probe_constants(_d(out))
    H = hermite_cardinals()
    # The cardinal functions: values and slopes at the ends.
    for nm, (v_l, v_r, d_l, d_r) in {"wl0": (1, 0, 0, 0), "wr0": (0, 1, 0, 0), "wl1": (0, 0, 1, 0),
                                     "wr1": (0, 0, 0, 1)}.items():
        u = H[nm]
        assert (u(-1.0), u(1.0)) == (v_l, v_r)
        h = 1e-6
        assert abs((u(-1 + h) - u(-1 - h)) / (2 * h) - d_l) < 1e-9
        assert abs((u(1 + h) - u(1 - h)) / (2 * h) - d_r) < 1e-9
    s, w = lobatto_interior()
    ip = lambda p, q: (w * H[p](s) * H[q](s)).sum() / (NS - 2)
    want = [ip("wl0", "wl1"), ip("wl1", "wl1"), ip("wl0", "wr1"), ip("wl1", "wr1"), ip("wr0", "wr1"),
            ip("wr1", "wr1"), ip("wl1", "wr0")]
    assert np.allclose(out, want, rtol=1e-13, atol=0)


def slope_lstsq(P, x, fixed):
    """The edge slopes as the discrete least-squares problem
    min sum_i <f - p_i, f - p_i>_i / h_i over the free slopes, p_i the cubic
    Hermite interpolant of interval i (edge values f(x_i), slopes s_i),
    <.,.>_i the interior Chebyshev-Lobatto quadrature on [x_i, x_i+1], h_i
    the half-width; with fixed = (fpa, fpb) the end slopes are given."""
    H = hermite_cardinals()
    s, w = lobatto_interior()
    n = len(x) - 1
    xd = x.astype(np.float64)
    fx = P.f(xd)
    rows, rhs = [], []
    for i in range(n):
        mid = 0.5 * np.float64(np.float32(x[i] + x[i + 1]))
        half = 0.5 * np.float64(np.float32(x[i + 1] - x[i]))
        fv = P.f(mid + half * s)
        sw = np.sqrt(w / (NS - 2) / half)
        base = fx[i] * H["wl0"](s) + fx[i + 1] * H["wr0"](s)
        A = np.zeros((len(s), n + 1))
        A[:, i] = half * H["wl1"](s)
        A[:, i + 1] = half * H["wr1"](s)
        rows.append(A * sw[:, None])
        rhs.append((fv - base) * sw)
    A = np.vstack(rows)
    y = np.concatenate(rhs)
    if fixed is None:
        sol, *_ = np.linalg.lstsq(A, y, rcond=None)
        return sol
    y = y - A[:, 0] * fixed[0] - A[:, n] * fixed[1]
    sol, *_ = np.linalg.lstsq(A[:, 1:n], y, rcond=None)
    return np.concatenate([[fixed[0]], sol, [fixed[1]]])


@pytest.mark.parametrize("name", ["lj", "morse", "morse3", "angle", "dihedral"])
@pytest.mark.parametrize("n", [10, 37])
def test_edge_slopes_solve_least_squares(probe, name, n):
    """potential_getfp_fixend / potential_getfp: the builder's slopes equal
    the least-squares solution assembled directly, and they solve the
    tridiagonal normal equations assembled from the quadrature (not the
    builder's constants) to a residual of 1e-12 of its scale."""
    P = Probe(probe, name)
    B = P.build()
    x = P.edges(B["al"], B["a"], B["b"], n)
    sl = P.slopes(x)
    free = P.free_ends()
    fixed = None if free else (float(P.fp(np.array([float(x[0])]))[0]), float(P.fp(np.array([float(x[n])]))[0]))
    ref = slope_lstsq(P, x, fixed)
    scale = np.abs(ref).max()
    assert np.abs(sl - ref).max() <= 1e-9 * scale, np.abs(sl - ref).max() / scale
    # Tridiagonal normal equations from the quadrature.
    H = hermite_cardinals()
    s, w = lobatto_interior()
    ip = lambda p, q: (w * H[p](s) * H[q](s)).sum() / (NS - 2)
    xd = x.astype(np.float64)
    fx = P.f(xd)
    hx = (x[1:] - x[:-1]).astype(np.float64)
    T = np.zeros((n + 1, n + 1))
    r = np.zeros(n + 1)

    def proj(i, card):
        mid = 0.5 * np.float64(np.float32(x[i] + x[i + 1]))
        half = 0.5 * np.float64(np.float32(x[i + 1] - x[i]))
        return (w * P.f(mid + half * s) * H[card](s)).sum() / (NS - 2)

    for i in range(n + 1):
        if i > 0:  # interval i - 1, where edge i is the right end
            T[i, i - 1] += ip("wl1", "wr1") * hx[i - 1]
            T[i, i] += ip("wr1", "wr1") * hx[i - 1]
            r[i] += 2 * (proj(i - 1, "wr1") - fx[i - 1] * ip("wl0", "wr1") - fx[i] * ip("wr0", "wr1"))
        if i < n:  # interval i, where edge i is the left end
            T[i, i] += ip("wl1", "wl1") * hx[i]
            T[i, i + 1] += ip("wl1", "wr1") * hx[i]
            r[i] += 2 * (proj(i, "wl1") - fx[i] * ip("wl0", "wl1") - fx[i + 1] * ip("wl1", "wr0"))
    if not free:
        T[0] = 0
        T[0, 0] = 1
        r[0] = fixed[0]
        T[n] = 0
        T[n, n] = 1
        r[n] = fixed[1]
    res = T @ sl - r
    assert np.abs(res).max() <= 1e-12 * (np.abs(T).max() * scale + np.abs(r).max())


# ---------------------------------------------------------------- interval fits


def phi_reference(fs_mid, fa, fb):
    """potential_getcoeffs's phi_j, j < 7: 2/N [(fa + (-1)^j fb) / 2 +
    sum_{k=1}^{N-1} f(m + h cos(k pi / N)) cos(j k pi / N)], with the cosines
    rounded to float as the reference stores them.  fa is the left edge's
    value, although cos(N pi / N) = -1 belongs to it only for even j."""
    k = np.arange(1, NS)
    phi = np.zeros(7)
    for j in range(7):
        cj = np.cos(j * k * np.pi / NS).astype(np.float32).astype(np.float64)
        phi[j] = ((fa + (1 - 2 * (j % 2)) * fb) / 2 + (fs_mid * cj).sum()) * 2.0 / NS
    return phi


def expected_chunk(fa, fb, dfa, dfb, phi):
    """The Hermite cubic with end values fa, fb and end slopes dfa, dfb (in
    s on [-1, 1]) in Chebyshev form (numpy), the least-squares corrections
    along the bubbles -3 + 4 T2 - T4 and -2 T1 + 3 T3 - T5 (both vanish with
    their slope at s = +-1) that bring the coefficients (first one doubled,
    as the reference stores it) closest to phi, and the monomials."""
    M = np.array([[1, -1, 1, -1], [1, 1, 1, 1], [0, 1, -2, 3], [0, 1, 2, 3]], dtype=np.float64)
    mono = np.linalg.solve(M, np.array([fa, fb, dfa, dfb]))
    cee = np.zeros(6)
    ch = npcheb.poly2cheb(mono)
    cee[:len(ch)] = ch
    cee[0] *= 2
    d_even = np.array([-6.0, 0, 4, 0, -1, 0])
    d_odd = np.array([0.0, -2, 0, 3, 0, -1])
    for d, idx in ((d_even, [0, 2, 4]), (d_odd, [1, 3, 5])):
        w = -np.dot(d[idx], cee[idx] - phi[idx]) / np.dot(d[idx], d[idx])
        cee = cee + w * d
    c = cee.copy()
    c[0] /= 2
    mono5 = npcheb.cheb2poly(c)
    mono5 = np.concatenate([mono5, np.zeros(6 - len(mono5))])
    return mono5[::-1], cee  # highest power first


@pytest.mark.parametrize("name", ["lj", "morse", "morse3", "harmonic", "angle", "dihedral"])
def test_interval_fits(probe, name):
    """Each interval of the chosen table: the quintic takes the edge values
    and slope * half-width at s = -1, 1; its coefficients are the Hermite
    cubic plus the degree-4/5 corrections toward phi, converted to
    monomials; the builder's error estimate is recomputed from the stored
    floats."""
    P = Probe(probe, name)
    B = P.build()
    n = B["n"]
    x = P.edges(B["al"], B["a"], B["b"], n)
    c, err = P.fit(x)
    # The chosen table's chunks are these fits.
    assert np.array_equal(c.view(np.uint32), B["c"].reshape(n + 1, 8)[1:].view(np.uint32))
    sl = P.slopes(x)
    fx = P.f(x.astype(np.float64))
    k = np.arange(1, NS)
    for i in range(n):
        mid = np.float64(np.float32(x[i] + x[i + 1]) / np.float32(2))
        half = np.float64(np.float32(x[i + 1] - x[i]) / np.float32(2))
        fa, fb = fx[i], fx[i + 1]
        dfa, dfb = sl[i] * half, sl[i + 1] * half
        scale = max(abs(fa), abs(fb), abs(dfa), abs(dfb), 1e-30)
        ci = c[i].astype(np.float64)
        assert ci[0] == np.float32(mid) and ci[1] == np.float32(1.0 / half)
        poly = ci[2:]  # s^5 .. s^0
        val = lambda s: np.polyval(poly, s)
        der = lambda s: np.polyval(np.polyder(poly), s)
        tol = 4e-7 * np.abs(poly).sum() + 1e-12 * scale
        assert abs(val(-1.0) - fa) <= tol and abs(val(1.0) - fb) <= tol, (i, val(-1.0) - fa, val(1.0) - fb)
        assert abs(der(-1.0) - dfa) <= 6 * tol and abs(der(1.0) - dfb) <= 6 * tol, i
        phi = phi_reference(P.f(mid + half * np.cos(k * np.pi / NS)), fa, fb)
        want, cee = expected_chunk(fa, fb, dfa, dfb, phi)
        wtol = 4e-7 * np.abs(want).sum() + 1e-12 * scale
        assert np.all(np.abs(poly - want) <= wtol), (i, poly - want)
        # Degree-4 and degree-5 terms are 8 cee4 and 16 cee5.
        assert abs(poly[1] - 8 * cee[4]) <= wtol and abs(poly[0] - 16 * cee[5]) <= wtol
    assert err == np.float32(builder_error(P, c, x)) or abs(err - builder_error(P, c, x)) <= 1e-6 * max(err, 1e-9)


def test_odd_phi_endpoint_order_does_not_change_the_fit():
    """The reference's odd-j phi differ from the Chebyshev-Lobatto
    coefficients of f: for f(s) = s the true phi_1 is 1 and the reference's
    is 1 - 4/N (its endpoints enter with the opposite sign).  The shift is the
    same for j = 1, 3, 5, and the degree-5 correction weighs them by
    (2, -3, -1) / 14, which sums to zero: both orders give the same chunk."""
    s = np.cos(np.arange(1, NS) * np.pi / NS)
    phi = phi_reference(s, -1.0, 1.0)
    assert abs(phi[1] - (1 - 4.0 / NS)) < 1e-6
    true1 = (2.0 / NS) * ((1.0 + 1.0) / 2 + (s * s).sum())
    assert abs(true1 - 1.0) < 1e-12
    rng = np.random.default_rng(3)
    for _ in range(20):
        fa, fb, dfa, dfb = rng.normal(size=4)
        fs = rng.normal(size=NS - 1)
        ref = phi_reference(fs, fa, fb)
        swapped = phi_reference(fs, fb, fa)  # (fb + (-1)^j fa) / 2: the Lobatto order
        assert np.allclose(ref[0::2], swapped[0::2], rtol=0, atol=1e-15)
        assert np.allclose(expected_chunk(fa, fb, dfa, dfb, ref)[0], expected_chunk(fa, fb, dfa, dfb, swapped)[0],
                           rtol=0, atol=1e-12)


# ---------------------------------------------------------------- tolerance, n


@pytest.mark.parametrize("name", sorted(CONFIGS))
def test_tolerance_met_and_n_minimal(probe, name):
    """The chosen table meets tol at the builder's nodes, measured
    independently from its stored coefficients (<= tol, no slack); with one
    interval fewer the same measure exceeds tol, so n is the smallest count
    that meets it (the doubling-then-bisection search assumes the error
    falls with n; every count it tried agrees)."""
    P = Probe(probe, name)
    B = P.build()
    n, tol = B["n"], P.tol
    x = P.edges(B["al"], B["a"], B["b"], n)
    c = B["c"].reshape(n + 1, 8)[1:]
    e_n = builder_error(P, c, x)
    assert e_n <= tol, (e_n, tol)
    tried = dict(B["tried"])
    assert tried[n] < tol
    for k, e in B["tried"]:
        assert (e < tol) == (k >= n) or (k < n and e == tol), (k, e)
    if n > 1:
        x1 = P.edges(B["al"], B["a"], B["b"], n - 1)
        c1, err1 = P.fit(x1)
        e_m = builder_error(P, c1, x1)
        assert abs(e_m - err1) <= 1e-6 * err1
        assert err1 >= tol and e_m >= tol * (1 - 1e-6), (n, err1, e_m)
    else:
        assert B["tried"][0] == (1, B["tried"][0][1])


@pytest.mark.parametrize("name", ["lj", "morse", "morse3"])
def test_tolerance_between_nodes(probe, name):
    """Between the builder's nodes (dense samples of every interval, double
    evaluation of the float coefficients) the error stays within 1% of tol:
    the 99 nodes per interval resolve the interpolation error."""
    P = Probe(probe, name)
    B = P.build()
    n, tol = B["n"], P.tol
    c = B["c"].reshape(n + 1, 8)[1:]
    worst = 0.0
    s = np.linspace(-1, 1, 2001)
    for i in range(n):
        mid, half = float(c[i, 0]), 1.0 / float(c[i, 1])
        fs = P.f(mid + half * s)
        e = np.abs(fs - horner_monomial(c[i], s)).max() / max(np.abs(fs).max(), 1.0)
        worst = max(worst, e)
    assert worst <= 1.01 * tol, worst


# ---------------------------------------------------------------- continuation


@pytest.mark.parametrize("name", ["lj", "morse", "morse3", "ljmol"])
def test_first_interval_is_quadratic_taylor(probe, name):
    """Interval 0 (r below a, x = (r - a) / a) is the quadratic Taylor
    expansion about interval 1's left edge of interval 1's polynomial without
    its degree-5 term: the reference's loop (MxPotential.cpp:1911-1919) runs
    over the five coefficients below it.  Value, slope and curvature at r = a
    equal that truncated polynomial's, and there are no terms above x^2; the
    value differs from interval 1's own by its s^5 coefficient."""
    P = Probe(probe, name)
    B = P.build()
    c = B["c"].reshape(-1, 8)
    a = float(B["a"])
    c0, c1 = c[0].astype(np.float64), c[1].astype(np.float64)
    assert c0[0] == np.float32(a) and c0[1] == np.float32(1.0 / a)
    assert np.all(c0[2:5] == 0.0)
    # interval 1 in r: s = (r - m) / h
    m1, ih1 = c1[0], c1[1]
    p1 = np.poly1d(np.concatenate([[0.0], c1[3:]]))  # the s^5 term dropped
    s_a = (a - m1) * ih1
    v1, d1, dd1 = p1(s_a), p1.deriv()(s_a) * ih1, p1.deriv(2)(s_a) * ih1 ** 2
    ih0 = c0[1]
    p0 = np.poly1d(c0[2:])
    x_a = (a - c0[0]) * ih0
    v0, d0, dd0 = p0(x_a), p0.deriv()(x_a) * ih0, p0.deriv(2)(x_a) * ih0 ** 2
    sc = np.abs(c1[2:]).sum()
    # interval 1's left edge is s = -1 up to the float rounding of its chunk
    assert abs(s_a + 1.0) <= 1e-5
    assert abs((v1 + c1[2] * s_a ** 5) - np.poly1d(c1[2:])(s_a)) <= 1e-9 * sc
    assert abs(v0 - v1) <= 1e-6 * sc
    assert abs(d0 - d1) <= 1e-5 * sc * ih1
    assert abs(dd0 - dd1) <= 1e-4 * sc * ih1 ** 2
