"""GPU parity: the HIP force path (through the C ABI) against the FP64 oracle.

Metrics (mixed-precision bar: force RMS relative error <= 2e-6):
  force_rms_rel = sqrt(mean |g - g_ref|^2) / sqrt(mean |g_ref|^2)
Energies are compared per term, the virial against max |W_ref|.  The all-FP64
instantiation of the same kernels must agree to ~1e-9.  Fixed-point
accumulation makes repeated evaluations bitwise identical.
"""
import numpy as np
import pytest

import golden_util
from apocharmm_amd import _lib, systems

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

MIXED_FORCE_RTOL = 2e-6
# the FP64 instantiation differs from the oracle only by the 32-bit
# fixed-point coordinate grid (box/2^32 ~ 1e-8 A), a ~1e-8 force effect
FP64_FORCE_RTOL = 5e-8
# energies: per-term error against the scale of all terms (the potential is a
# small sum of large, cancelling terms)
# (FP32 pair energies carry ~1e-7 relative rounding each; their sum over
# ~1e6 pairs lands near 1e-7 of the scale)
MIXED_ENERGY_RTOL = 2e-7
FP64_ENERGY_RTOL = 1e-10


def energy_scale(ref):
    return sum(abs(v) for k, v in ref["energy"].items() if k != "epot")


def _fm(spec, precision="mixed"):
    from apocharmm_amd.force_manager import ForceManager
    return ForceManager(spec, precision).initialize()


def gpu_eval(spec, precision="mixed", terms=0, xyz=None, fm=None):
    fm = fm or _fm(spec, precision)
    xyzq = torch.tensor(systems.xyzq_of(spec, xyz), device="cuda")
    fm.calcForce(xyzq, reset=True, calcEnergy=True, calcVirial=True, terms=terms)
    torch.cuda.synchronize()
    g = fm.getForces().cpu().numpy().copy()
    gi = fm.getForcesFixed().cpu().numpy().copy()
    return dict(grad=g, grad_fixed=gi, energy=fm.getEnergyComponents(), virial=fm.getVirial(),
                fm=fm)


ORACLE_TERMS = {
    0: ("bonded", "direct", "pairs14", "recip", "self"),
    _lib.APO_TERM_DIRECT: ("direct",),
    _lib.APO_TERM_RECIP: ("recip",),
    _lib.APO_TERM_BONDED: ("bonded",),
    _lib.APO_TERM_PAIRS14: ("pairs14",),
}


def oracle_eval(oracle, spec, terms=0):
    s = oracle.prepare(spec)
    return oracle.evaluate(s, systems.xyzq_of(spec), terms=ORACLE_TERMS[terms])


def force_rms_rel(g, ref):
    return float(np.sqrt(np.mean(np.sum((g - ref) ** 2, 1))) / np.sqrt(np.mean(np.sum(ref ** 2, 1))))


def report(tag, got, ref):
    err = force_rms_rel(got["grad"], ref["grad"])
    print("\n[%s] force_rms_rel=%.3e" % (tag, err), flush=True)
    for k in ("bond", "angle", "dihe", "ewks", "ewse", "ewex", "elec", "vdw", "epot"):
        r = ref["energy"][k]
        if r != 0.0:
            print("   %-5s gpu %.10g oracle %.10g rel %.2e" % (k, got["energy"][k], r,
                                                              abs(got["energy"][k] - r) / abs(r)))
    return err


SMALL = {
    "c1_water3000": lambda: systems.config1(),
    "chain300_w900": lambda: systems.solvated_chain(n_chain=300, n_waters=900, L=32.0, fft_grid=32),
}


@pytest.mark.parametrize("name", list(SMALL))
def test_mixed_total_matches_oracle(oracle, name):
    spec = SMALL[name]()
    print("case", name, "natom", len(spec["charges"]), flush=True)
    got = gpu_eval(spec)
    ref = oracle_eval(oracle, spec)
    err = report(name + " mixed", got, ref)
    assert err <= MIXED_FORCE_RTOL
    scale = energy_scale(ref)
    for k in ("bond", "angle", "dihe", "ewks", "ewse", "ewex", "elec", "vdw", "epot"):
        r = ref["energy"][k]
        assert abs(got["energy"][k] - r) <= MIXED_ENERGY_RTOL * scale, k
    wmax = np.abs(ref["virial"]).max()
    assert np.abs(got["virial"] - ref["virial"]).max() <= 1e-5 * wmax


@pytest.mark.parametrize("term", [_lib.APO_TERM_DIRECT, _lib.APO_TERM_RECIP, _lib.APO_TERM_BONDED,
                                  _lib.APO_TERM_PAIRS14])
def test_mixed_terms_match_oracle(oracle, term):
    spec = SMALL["chain300_w900"]()
    print("case chain300 term", hex(term), flush=True)
    got = gpu_eval(spec, terms=term)
    ref = oracle_eval(oracle, spec, terms=term)
    total = oracle_eval(oracle, spec)
    report("term %x" % term, got, ref)
    # each term's force error measured against the total force of the system
    dg = got["grad"] - ref["grad"]
    err = np.sqrt(np.mean(np.sum(dg ** 2, 1))) / np.sqrt(np.mean(np.sum(total["grad"] ** 2, 1)))
    print("   term error / total force rms = %.3e" % err, flush=True)
    assert err <= 1.2e-6


@pytest.mark.parametrize("name", list(SMALL))
def test_fp64_instantiation_matches_oracle(oracle, name):
    spec = SMALL[name]()
    print("case", name, "fp64", flush=True)
    got = gpu_eval(spec, precision="fp64")
    ref = oracle_eval(oracle, spec)
    err = report(name + " fp64", got, ref)
    assert err <= FP64_FORCE_RTOL
    scale = energy_scale(ref)
    for k in ("ewks", "ewex", "elec", "vdw", "epot"):
        r = ref["energy"][k]
        assert abs(got["energy"][k] - r) <= FP64_ENERGY_RTOL * scale, k


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
@pytest.mark.parametrize("name", golden_util.NAMES)
def test_gpu_matches_c47b1_golden(name, precision):
    """The reference's own tolerances against CHARMM c47b1
    (test/unittests/unittest-eneForces.cpp:110-365), for both precisions.  The
    water dimer's elec energy (0.122 kcal/mol) is a sum of 9 pair terms of
    ~30 kcal/mol: the mixed path meets the 1e-5 margin because energy steps
    evaluate pair energies in FP64 from the fixed-point geometry."""
    spec, xyzq, ref = golden_util.load(name)
    spec = dict(spec)
    spec.update(xyz=xyzq[:, :3], masses=np.ones(len(xyzq)), name=name)
    systems._empty_lists(spec)
    print("case golden", name, precision, flush=True)
    got = gpu_eval(spec, precision=precision)
    for key, val in ref["energy"].items():
        tol = max(ref["energy_atol"], ref["energy_rtol"] * abs(val))
        assert abs(got["energy"][key] - val) <= tol, (key, got["energy"][key], val)
    assert np.abs(got["grad"] - ref["grad"]).max() <= ref["force_atol"]
    assert np.abs(got["virial"].reshape(-1) - ref["virial"]).max() <= ref["virial_atol"]


def test_bitwise_reproducible():
    spec = systems.config1()
    print("case c1 determinism", flush=True)
    a = gpu_eval(spec)
    fm = a["fm"]
    b = gpu_eval(spec, fm=fm)
    c = gpu_eval(spec)  # fresh list build
    for other in (b, c):
        np.testing.assert_array_equal(a["grad_fixed"], other["grad_fixed"])
        assert a["energy"] == other["energy"]
        np.testing.assert_array_equal(a["virial"], other["virial"])


def brute_pairs_within(spec, rc):
    x = spec["xyz"]
    L = spec["box"]
    n = len(x)
    excl = set()
    from oracle import oracle as O
    ptr, idx, _, _ = O.exclusion_lists(n, spec["bonds"])
    for i in range(n):
        for j in idx[ptr[i]:ptr[i + 1]]:
            if j > i:
                excl.add((i, int(j)))
    cnt = 0
    for i in range(n - 1):
        d = x[i + 1:] - x[i]
        d -= L * np.round(d / L)
        r2 = np.sum(d * d, 1)
        js = np.nonzero(r2 < rc * rc)[0] + i + 1
        cnt += len(js) - sum((i, int(j)) in excl for j in js)
    return cnt


def test_list_covers_every_pair_inside_list_cutoff():
    """The reference's debug cross-check (src/CudaNeighborListBuild.cu:2215-2351):
    pairs covered by the tiles == brute-force minimum-image count inside cutnb."""
    spec = systems.config1()
    got = gpu_eval(spec)
    st = got["fm"].listStats()
    print("list stats", st, flush=True)
    want = brute_pairs_within(spec, spec["cutnb"])
    # float vs double distance near the cutoff can move a handful of pairs
    assert abs(st["pairs_in_list"] - want) <= max(3, want * 1e-6), (st["pairs_in_list"], want)


def test_small_box_below_list_cutoff():
    """500 waters: L = 24.6 A < 2 * cutnb; minimum image holds for r_off."""
    spec = systems.batched_system(0, 500)
    print("case small box L=%.2f" % spec["box"][0], flush=True)
    got = gpu_eval(spec)
    from oracle import oracle as O
    ref = oracle_eval(O, spec)
    assert force_rms_rel(got["grad"], ref["grad"]) <= MIXED_FORCE_RTOL


@pytest.mark.parametrize("n_waters", [1200, 17146])
def test_odd_pme_grid_against_oracle(oracle, n_waters):
    """Boxes whose ~1 A PME grid is odd (27^3 = 3^9 points, 81^3 = four
    radix-3 passes): odd grids have no Nyquist bin and the x/L + 1/2 origin is
    not an integer grid shift, so the spread, FFT and gather must follow the
    oracle's convention exactly."""
    L = (n_waters / systems.WATER_DENSITY) ** (1.0 / 3.0)
    K = systems.smooth_grid(L)
    assert K % 2 == 1
    spec = systems.tip3p_box(n_waters, L, K, seed=6000 + n_waters)
    print("case odd grid %d waters K=%d" % (n_waters, K), flush=True)
    got = gpu_eval(spec)
    ref = oracle_eval(oracle, spec)
    err = report("odd K=%d" % K, got, ref)
    assert err <= MIXED_FORCE_RTOL
    scale = energy_scale(ref)
    for k in ("ewks", "elec", "vdw", "ewex"):
        assert abs(got["energy"][k] - ref["energy"][k]) <= MIXED_ENERGY_RTOL * scale, k


@pytest.mark.parametrize("n_waters,K", [(700, 28), (1200, 33), (600, 26)])
def test_large_radix_pme_grid_against_oracle(oracle, n_waters, K):
    """Grids whose FFT plans need radix 7, 11 or 13 (28 = 4 x 7, 33 = 3 x 11,
    26 = 2 x 13): the FFT kernels' variant that holds those stages (the other
    variant, for plans of radices up to 5, runs every other test)."""
    L = (n_waters / systems.WATER_DENSITY) ** (1.0 / 3.0)
    spec = systems.tip3p_box(n_waters, L, K, seed=6100 + K)
    print("case radix grid %d waters L=%.2f K=%d" % (n_waters, L, K), flush=True)
    got = gpu_eval(spec)
    ref = oracle_eval(oracle, spec)
    err = report("radix K=%d" % K, got, ref)
    assert err <= MIXED_FORCE_RTOL
    scale = energy_scale(ref)
    for k in ("ewks", "elec", "vdw", "ewex"):
        assert abs(got["energy"][k] - ref["energy"][k]) <= MIXED_ENERGY_RTOL * scale, k


def test_config2_full_size_against_oracle(oracle):
    """98,304-atom TIP3P box, 100^3 grid: the benchmark workload at step 0."""
    spec = systems.config2()
    print("case c2 98304 atoms", flush=True)
    got = gpu_eval(spec)
    ref = oracle_eval(oracle, spec)
    err = report("c2 mixed", got, ref)
    assert err <= MIXED_FORCE_RTOL
    scale = energy_scale(ref)
    for k in ("bond", "angle", "ewks", "ewse", "ewex", "elec", "vdw", "epot"):
        assert abs(got["energy"][k] - ref["energy"][k]) <= MIXED_ENERGY_RTOL * scale, k
    wmax = np.abs(ref["virial"]).max()
    print("   virial max |dW| / max |W| = %.2e" % (np.abs(got["virial"] - ref["virial"]).max()
                                                   / wmax), flush=True)
    assert np.abs(got["virial"] - ref["virial"]).max() <= 1e-5 * wmax
    # size-independent invariant: the pair, exclusion and bonded gradients sum
    # to zero; the PME gradient sum is the small discretisation residue
    g = got["grad_fixed"].astype(np.float64) / 2.0 ** 40
    assert np.abs(g.sum(0)).max() < 1e-2 * np.sqrt(len(g))


def test_config3_full_size_against_oracle(oracle):
    """The 23,558-atom solvated chain (2,489-atom chain with 2,491 LJ types,
    bonds, angles, dihedrals and 1-4 pairs, in 7,023 waters, 64^3 grid) at
    step 0: every force term and energy against the oracle."""
    spec = systems.config3()
    print("case c3 %d atoms, %d LJ types" % (len(spec["charges"]), len(spec["eps"])), flush=True)
    got = gpu_eval(spec)
    ref = oracle_eval(oracle, spec)
    err = report("c3 mixed", got, ref)
    assert err <= MIXED_FORCE_RTOL
    scale = energy_scale(ref)
    for k in ("bond", "angle", "dihe", "ewks", "ewse", "ewex", "elec", "vdw", "epot"):
        assert abs(got["energy"][k] - ref["energy"][k]) <= MIXED_ENERGY_RTOL * scale, k
    wmax = np.abs(ref["virial"]).max()
    assert np.abs(got["virial"] - ref["virial"]).max() <= 1e-5 * wmax


def test_calc_force_rejects_unknown_flag_bits():
    """apo_nb_calc_force accepts the APO_CALC_*, APO_RESET_LIST and APO_TERM_*
    bits only (the others are reserved for the library's own step path)."""
    from apocharmm_amd.force_manager import ForceManager
    sp = systems.config1()
    fm = ForceManager(sp).initialize()
    x = torch.tensor(systems.xyzq_of(sp), device="cuda")
    fm.calcForce(x, reset=True)
    for bit in (4, 7, 16, 26, 27, 28, 30):
        with pytest.raises(_lib.ApoInvalidArgument):
            fm.calcForce(x, terms=1 << bit)
    fm.calcForce(x)  # the system stays usable
    torch.cuda.synchronize()
    assert np.all(np.isfinite(fm.getForces().cpu().numpy()))
    fm.close()
