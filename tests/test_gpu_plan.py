"""GPU parity of the batched plan (tga_plan_* / tga_search, the level-3 replacement of
acquire_all_*) against the oracle run per (antenna, PRN) with the reference's semantics:
de-mean over the whole record, truncate to N = min(2P, record), code_spectra of the N-sample
replica, correlate_code (acquisition.rs:281-317, correlate.rs:91-183).  Covers the fused
P = 16368 kernels, the generic folded path, the generic unfolded path (record < 2P), 1- and
2-bit input, several systems sharing one group, and the plan's error paths."""
import numpy as np
import pytest

import tga
from tga import synth
from oracle import codes as ocodes
from parity import assert_peak_matches, near_ties, oracle_search

pytestmark = pytest.mark.gpu


def run_case(fs, L, n_ant, fmt, systems, signals, seed):
    """systems: list of (name, prns, fc list). Returns (plan, peaks, packed)."""
    packed = synth.synth_observation(n_ant, L, fs, signals, fmt=fmt, seed=seed)
    plan = tga.Plan(fs, n_ant, L, [tga.SystemSearch(n, fc, prns) for n, prns, fc in systems],
                    input_format=fmt)
    peaks = plan.search(packed)
    return plan, peaks, packed


def check_against_oracle(plan, peaks, packed, fs, L, fmt, systems, ants=None, prn_idx=None):
    levels = synth.unpack_levels(packed, L, fmt)
    n_checked = 0
    for si, (name, prns, fc) in enumerate(systems):
        P = ocodes.samples_per_period(name, fs)
        N = min(2 * P, L)
        for a in (ants if ants is not None else range(plan.n_ant)):
            for pi, prn in enumerate(plan.prns[si]):
                if prn_idx is not None and pi not in prn_idx:
                    continue
                code = ocodes.code_resampled(name, prn, float(P), N)
                ref, sig, pp, cc, ca = oracle_search(levels[a], fs, P, N, fc, code)
                ties = near_ties(sig, pp, cc, ca, fc, N, P)
                assert_peak_matches(peaks[si][a, pi], ref, P, ties,
                                    what=f"{name} ant {a} prn {prn}")
                n_checked += 1
    return n_checked


def grid(center, band, step):
    n = int(round(2 * band / step)) + 1
    return [center - band + i * step for i in range(n)]


FS = 16.368e6
IF = 4.092e6


@pytest.fixture(params=[0, 1], ids=["valu31", "mfma31"])
def fused_variant(request, monkeypatch):
    """Both radix-31 implementations of the fused rows kernel (the autotuner picks one)."""
    monkeypatch.setenv("TGA_FUSED_VARIANT", str(request.param))
    return request.param


@pytest.fixture(params=[0, 1, 2, 3, 5],
                ids=["split-valu31", "split-mfma31", "split-subrows", "split-fourstep",
                     "split-half"])
def split_variant(request, monkeypatch):
    """Every FP32 form of the long rows: R0 sub-rows of 16368 with either radix-31
    implementation (k_rows), the sub-row transform that keeps each lane's best output pair
    (k_subrows), 2 R0 half sub-rows of 8184 points (k_half_pre + k_halfrows), and the
    four-step N1 x 496 form."""
    monkeypatch.setenv("TGA_SPLIT_VARIANT", str(request.param))
    return request.param


def split_path(variant, r0):
    """The describe() path of an FP32 split group of R0 x 16368-sample periods."""
    if variant == 3:
        return "split132x496" if r0 == 4 else "split330x496"
    return f"split{r0}x16368"


@pytest.mark.parametrize("fmt", [1, 2])
def test_fused_ca_plan_matches_oracle(gpu, fmt, fused_variant):
    L = 163680
    signals = [dict(system="gps", prn=5, cn0=52.0, doppler=1000.0, delay=1234, phase=0.3),
               dict(system="gps", prn=12, cn0=45.0, doppler=-2500.0, delay=9000, phase=1.1),
               dict(system="sbas", prn=131, cn0=50.0, doppler=500.0, delay=15000, phase=2.0)]
    fc = grid(IF, 3000.0, 500.0)
    systems = [("gps", [1, 5, 12, 30], fc), ("sbas", [120, 131], fc)]
    plan, peaks, packed = run_case(FS, L, 2, fmt, systems, signals, seed=11 + fmt)
    assert "fused16368" in plan.describe()
    assert f"rows_variant={fused_variant}" in plan.describe()
    check_against_oracle(plan, peaks, packed, FS, L, fmt, systems)
    # injected signals are found at the expected code phase and nearest Doppler bin
    for si, sv in ((0, 1), (0, 2), (1, 1)):
        name = systems[si][0]
        s = [x for x in signals if x["system"] == name and x["prn"] == plan.prns[si][sv]][0]
        for a in range(2):
            pk = peaks[si][a, sv]
            # 16 samples per chip: the correlation top is flat to within the noise at
            # 45 dB-Hz over 2 ms, so the weakest SV may land a few samples off
            tol = 0 if s["cn0"] >= 50 else 4
            d = (int(pk["best_codephase"]) - (16368 - s["delay"])) % 16368
            assert min(d, 16368 - d) <= tol
            assert abs(fc[int(pk["best_freq_idx"])] - IF - s["doppler"]) <= 250.0


def test_fused_galileo_at_4092khz(gpu, fused_variant):
    # E1-C at fs = 4.092 MHz has P = 16368 samples per 4 ms period: the fused kernels serve it.
    fs, L = 4.092e6, 2 * 16368 + 1000
    signals = [dict(system="galileo", prn=11, cn0=50.0, doppler=-700.0, delay=4321)]
    fc = grid(IF, 1000.0, 250.0)
    systems = [("galileo", [3, 11], fc)]
    plan, peaks, packed = run_case(fs, L, 1, 1, systems, signals, seed=5)
    assert "fused16368" in plan.describe()
    check_against_oracle(plan, peaks, packed, fs, L, 1, systems)
    assert int(peaks[0][0, 1]["best_codephase"]) == (16368 - 4321) % 16368


def test_generic_folded_plan_matches_oracle(gpu):
    fs, L = 2.046e6, 3 * 2046
    signals = [dict(system="gps", prn=3, cn0=55.0, doppler=600.0, delay=700),
               dict(system="qzss", prn=193, cn0=55.0, doppler=-300.0, delay=100)]
    fc = grid(IF, 1500.0, 300.0)
    systems = [("gps", [3, 4], fc), ("qzss", [193, 200], fc)]
    plan, peaks, packed = run_case(fs, L, 3, 2, systems, signals, seed=7)
    assert "generic-folded" in plan.describe()
    check_against_oracle(plan, peaks, packed, fs, L, 2, systems)


def test_generic_unfolded_short_record(gpu):
    # record shorter than two periods: N = L (acquisition.rs:287), no folding
    fs, L = 2.046e6, 3000
    signals = [dict(system="gps", prn=9, cn0=58.0, doppler=0.0, delay=222)]
    fc = grid(IF, 900.0, 300.0)
    systems = [("gps", [9, 10], fc)]
    plan, peaks, packed = run_case(fs, L, 2, 1, systems, signals, seed=3)
    assert "generic-unfolded" in plan.describe()
    check_against_oracle(plan, peaks, packed, fs, L, 1, systems)


def test_plan_default_prn_ranges_and_layout(gpu):
    fs, L = 2.046e6, 2 * 2046
    fc = [IF]
    plan = tga.Plan(fs, 1, L, [tga.SystemSearch("sbas", fc), tga.SystemSearch("qzss", fc)])
    assert list(plan.prns[0]) == list(range(120, 159))
    assert list(plan.prns[1]) == list(range(184, 207))
    assert plan.n_peaks == 39 + 23
    packed = synth.synth_observation(1, L, fs, [], fmt=1, seed=1)
    a = plan.search(packed)
    b = plan.search(packed)
    for x, y in zip(a, b):  # deterministic
        assert x.tobytes() == y.tobytes()


def test_plan_errors(gpu):
    fc = [IF]
    with pytest.raises(tga.TgaError) as e:
        tga.Plan(FS, 1, 163680, [tga.SystemSearch("gps", [])])
    assert e.value.name == "TGA_ERR_EMPTY_BINS"
    with pytest.raises(tga.TgaError) as e:
        tga.Plan(FS, 1, 163680, [tga.SystemSearch("gps", fc, [0, 1])])
    assert e.value.name == "TGA_ERR_PRN_RANGE"
    with pytest.raises(tga.TgaError) as e:
        tga.Plan(FS, 1, 163680, [tga.SystemSearch("gps", fc)], input_format=3)
    assert e.value.name == "TGA_ERR_INVALID_ARGUMENT"
    plan = tga.Plan(FS, 2, 163680, [tga.SystemSearch("gps", fc, [1])])
    with pytest.raises(tga.TgaError) as e:
        plan.search(np.zeros((3, plan.row_bytes), np.uint8))
    assert e.value.name == "TGA_ERR_ANTENNA_RANGE"


def test_masked_second_peak_many_code_phases(gpu, fused_variant):
    # A 60 dB-Hz signal at 128 code phases (one per antenna): the main lobe's neighbours within
    # cw = 16 samples are nearly as strong as the peak, so a top-2 set of a lane that held two
    # outputs closer than 2 cw would report one of them as the second peak (a grouping with such
    # pairs affects a few percent of the phases, so 128 phases find it with high probability).
    # Both radix-31 implementations must give the oracle's masked second peak at every phase.
    n_ant = 128
    L = 2 * 16368
    rows = []
    delays = [int(d) for d in np.random.default_rng(31).integers(0, 16368, n_ant)]
    for a, d in enumerate(delays):
        sig = [dict(system="gps", prn=4, cn0=60.0, doppler=500.0, delay=d)]
        rows.append(synth.synth_observation(1, L, FS, sig, fmt=1, seed=100 + a)[0])
    packed = np.stack(rows)
    fc = grid(IF + 500.0, 500.0, 500.0)
    systems = [("gps", [4], fc)]
    plan = tga.Plan(FS, n_ant, L, [tga.SystemSearch("gps", fc, [4])])
    assert "fused16368" in plan.describe()
    peaks = plan.search(packed)
    assert check_against_oracle(plan, peaks, packed, FS, L, 1, systems) == n_ant
    for a, d in enumerate(delays):
        assert int(peaks[0][a, 0]["best_codephase"]) % 16368 == (16368 - d) % 16368


def test_split_e1c_at_16368khz(gpu, split_variant):
    # E1-C at 16.368 MHz: P = 65472 = 4 x 16368, rows transformed as 4 twiddled sub-rows by the
    # fused LDS kernel; the second peak comes from recomputing the winning row
    fs, L = FS, 2 * 65472 + 500
    # Doppler on the grid: 8 ms of coherent integration has its first null 125 Hz off
    signals = [dict(system="galileo", prn=11, cn0=52.0, doppler=500.0, delay=40000),
               dict(system="galileo", prn=27, cn0=49.0, doppler=-1500.0, delay=123)]
    fc = grid(IF, 1500.0, 500.0)
    systems = [("galileo", [3, 11, 27], fc)]
    plan, peaks, packed = run_case(fs, L, 2, 1, systems, signals, seed=21)
    assert split_path(split_variant, 4) in plan.describe()
    assert f"rows_variant={split_variant}" in plan.describe()
    check_against_oracle(plan, peaks, packed, fs, L, 1, systems)
    assert int(peaks[0][0, 1]["best_codephase"]) == (65472 - 40000) % 65472
    assert int(peaks[0][1, 2]["best_codephase"]) == (65472 - 123) % 65472


def test_split_row_tiles_ragged(gpu, split_variant):
    # long rows are enumerated in tiles of 8 forward spectra x 8 replica spectra, one hypothesis
    # at a time: 10 PRNs and 2 x 5 (bin, h) spectra leave full and ragged tiles in both
    # directions, and every row must still land at its own index
    fs, L = FS, 2 * 65472
    signals = [dict(system="galileo", prn=9, cn0=52.0, doppler=-1000.0, delay=31000)]
    fc = grid(IF, 1000.0, 500.0)
    systems = [("galileo", list(range(1, 11)), fc)]
    plan, peaks, packed = run_case(fs, L, 1, 1, systems, signals, seed=23)
    assert split_path(split_variant, 4) in plan.describe()
    assert check_against_oracle(plan, peaks, packed, fs, L, 1, systems) == 10
    assert int(peaks[0][0, 8]["best_codephase"]) == (65472 - 31000) % 65472


def test_split_tile_shapes_give_identical_peaks(gpu, monkeypatch):
    # the tile shape only reorders the long rows (and which spectra share the caches): every
    # shape, square, skewed either way and ragged against 10 PRNs x 10 (bin, h) spectra, gives
    # bit-identical peaks, matching the oracle
    fs, L = FS, 2 * 65472
    signals = [dict(system="galileo", prn=4, cn0=52.0, doppler=500.0, delay=1234)]
    fc = grid(IF, 1000.0, 500.0)
    systems = [("galileo", list(range(1, 11)), fc)]
    packed = synth.synth_observation(2, L, fs, signals, fmt=1, seed=29)
    base = None
    # one radix-31 implementation throughout (the two round differently)
    monkeypatch.setenv("TGA_SPLIT_VARIANT", "1")
    for t in ("8x8", "2x32", "32x2", "3x7", "1x1"):
        monkeypatch.setenv("TGA_SPLIT_TILE", t)
        plan = tga.Plan(fs, 2, L, [tga.SystemSearch(n, f, p) for n, p, f in systems])
        assert f"tile={t}" in plan.describe()
        pk = plan.search(packed)[0]
        if base is None:
            base = pk
            assert check_against_oracle(plan, [pk], packed, fs, L, 1, systems) == 20
        for f in pk.dtype.names:
            assert np.array_equal(pk[f], base[f]), (t, f)
    # the four-step form on every tile shape: bit-identical among themselves, and the oracle's
    monkeypatch.setenv("TGA_SPLIT_VARIANT", "3")
    base4 = None
    for t in ("8x8", "2x32", "3x7"):
        monkeypatch.setenv("TGA_SPLIT_TILE", t)
        plan = tga.Plan(fs, 2, L, [tga.SystemSearch(n, f, p) for n, p, f in systems])
        assert f"tile={t}" in plan.describe() and "split132x496" in plan.describe()
        pk = plan.search(packed)[0]
        if base4 is None:
            base4 = pk
            assert check_against_oracle(plan, [pk], packed, fs, L, 1, systems) == 20
        for f in pk.dtype.names:
            assert np.array_equal(pk[f], base4[f]), (t, f)
    # unpinned, the autotuner times the four FP32 forms x 5 shapes
    monkeypatch.delenv("TGA_SPLIT_TILE")
    monkeypatch.delenv("TGA_SPLIT_VARIANT")
    desc = tga.Plan(fs, 2, L, [tga.SystemSearch(n, f, p) for n, p, f in systems]).describe()
    for c in ("v2t8x8:", "v2t32x2:", "v1t8x8:", "v1t4x16:", "v0t16x4:", "v0t2x32:", "v1t32x2:",
              "v3t8x8:", "v3t4x16:", "v5t8x8:", "v5t32x2:"):
        assert c in desc, desc
    assert "v3t2x32:" not in desc  # the four-step form is timed on two shapes


@pytest.mark.parametrize("system,prn", [("beidou", 19), ("l1c", 7)])
def test_split_b1c_l1c_at_16368khz(gpu, split_variant, system, prn):
    # B1C / L1C at 16.368 MHz: P = 163680 = 10 x 16368 (BOC(1,1) half-chips)
    fs, L = FS, 2 * 163680
    # Doppler on the grid: 20 ms of coherent integration has its first null 50 Hz off
    signals = [dict(system=system, prn=prn, cn0=50.0, doppler=-500.0, delay=99999)]
    fc = grid(IF, 1000.0, 500.0)
    systems = [(system, [prn, prn + 1], fc)]
    plan, peaks, packed = run_case(fs, L, 1, 2, systems, signals, seed=prn)
    assert split_path(split_variant, 10) in plan.describe()
    check_against_oracle(plan, peaks, packed, fs, L, 2, systems)
    assert int(peaks[0][0, 0]["best_codephase"]) == (163680 - 99999) % 163680


def check_nc_against_oracle(plan, peaks, packed, fs, L, fmt, systems, epochs, prn_idx=None):
    from oracle import correlate as ocor
    levels = synth.unpack_levels(packed, L, fmt)
    for si, (name, prns, fc) in enumerate(systems):
        P = ocodes.samples_per_period(name, fs)
        for a in range(plan.n_ant):
            lv = levels[a].astype(np.float64)
            sig = (lv - lv.sum() / len(lv)).astype(np.float32)
            pp = ocor.phasepoints(fs, epochs * P)
            for j, prn in enumerate(plan.prns[si]):
                if prn_idx is not None and j not in prn_idx:
                    continue
                code = ocodes.code_resampled(name, prn, float(P), P)
                ref = ocor.noncoherent(sig, pp, code, fc, P, epochs, fs)
                got = peaks[si][a, j]
                what = f"{name} {prn} ant {a}"
                assert abs(float(got["best_peak"]) - ref["best_peak"]) <= \
                    1e-5 * ref["best_peak"], what
                vals = sorted((v for _, v in ref["rows"]), reverse=True)
                if len(vals) > 1 and vals[1] >= vals[0] * (1 - 2e-5):
                    continue  # near-tie between bins
                assert int(got["best_freq_idx"]) == ref["best_freq_idx"], what
                assert int(got["best_codephase"]) == ref["best_codephase"], what
                assert abs(float(got["second_peak"]) - ref["second_peak"]) <= \
                    1e-5 * ref["second_peak"], what


@pytest.mark.parametrize("fmt", [1, 2])
def test_noncoherent_plan_matches_oracle(gpu, fmt):
    # 10 one-period epochs of a 10 ms record, |y|^2 summed (an extension of the reference)
    L, E = 163680, 10
    signals = [dict(system="gps", prn=6, cn0=41.0, doppler=1500.0, delay=7000),
               dict(system="sbas", prn=133, cn0=44.0, doppler=-2000.0, delay=300)]
    fc = grid(IF, 2500.0, 500.0)
    systems = [("gps", [6, 7], fc), ("sbas", [133], fc)]
    packed = synth.synth_observation(2, L, FS, signals, fmt=fmt, seed=40 + fmt)
    plan = tga.Plan(FS, 2, L, [tga.SystemSearch(n, f, p, nc_epochs=E) for n, p, f in systems],
                    input_format=fmt)
    assert "fused16368-noncoherent" in plan.describe()
    peaks = plan.search(packed)
    check_nc_against_oracle(plan, peaks, packed, FS, L, fmt, systems, E)
    for a in range(2):
        assert int(peaks[0][a, 0]["best_codephase"]) == (16368 - 7000) % 16368
        assert int(peaks[1][a, 0]["best_codephase"]) == (16368 - 300) % 16368


@pytest.mark.parametrize("groups", ["8", "16", "32"])
def test_noncoherent_groups_match(gpu, monkeypatch, groups):
    # every PRN grouping of the non-coherent rows kernel (the autotuner's candidates) computes
    # the same rows: bit-identical peaks to the plan's own choice, and the oracle on two PRNs
    L, E = 163680, 10
    signals = [dict(system="gps", prn=9, cn0=42.0, doppler=-1000.0, delay=5000)]
    fc = grid(IF, 1500.0, 500.0)
    prns = list(range(1, 21))
    packed = synth.synth_observation(1, L, FS, signals, fmt=1, seed=77)
    monkeypatch.delenv("TGA_ROWS_GROUPS", raising=False)
    base = tga.Plan(FS, 1, L, [tga.SystemSearch("gps", fc, prns, nc_epochs=E)]).search(packed)
    monkeypatch.setenv("TGA_ROWS_GROUPS", groups)
    plan = tga.Plan(FS, 1, L, [tga.SystemSearch("gps", fc, prns, nc_epochs=E)])
    assert f"groups={groups}" in plan.describe()
    peaks = plan.search(packed)
    for f in ("best_peak", "second_peak", "best_codephase", "best_freq_idx"):
        np.testing.assert_array_equal(peaks[0][f], base[0][f], err_msg=f)
    check_nc_against_oracle(plan, peaks, packed, FS, L, 1, [("gps", prns, fc)], E,
                            prn_idx=[8, 9])
    assert int(peaks[0][0, 8]["best_codephase"]) == (16368 - 5000) % 16368


def test_noncoherent_plan_errors(gpu):
    fc = [IF]
    with pytest.raises(tga.TgaError):  # more epochs than the record holds
        tga.Plan(FS, 1, 5 * 16368, [tga.SystemSearch("gps", fc, [1], nc_epochs=6)])
    with pytest.raises(tga.TgaError):  # no fused path for E1-C periods
        tga.Plan(FS, 1, 4 * 65472, [tga.SystemSearch("galileo", fc, [1], nc_epochs=2)])


def test_split_unfolded_10ms_b1c(gpu, split_variant):
    # a 10 ms record holds one B1C period: N = P = 163680, unfolded rows through the split path,
    # only the [c,c] rows transformed (correlate.rs:53-56: the hypotheses coincide)
    fs, L = FS, 163680
    signals = [dict(system="beidou", prn=30, cn0=55.0, doppler=1000.0, delay=12345)]
    fc = grid(IF, 1000.0, 500.0)
    systems = [("beidou", [30, 31], fc)]
    plan, peaks, packed = run_case(fs, L, 2, 1, systems, signals, seed=12)
    assert split_path(split_variant, 10) in plan.describe()
    check_against_oracle(plan, peaks, packed, fs, L, 1, systems)
    assert int(peaks[0][0, 0]["best_codephase"]) == (163680 - 12345) % 163680


@pytest.mark.parametrize("system", ["gps", "galileo", "beidou"])
def test_constant_observation_every_power_ties(gpu, split_variant, system):
    # A constant record is all zeros after the whole-record de-mean, so every correlation power
    # is exactly 0 on both sides: the reference's last-maximum rules then fix the answer (last
    # sample of the row, last bin, [c,c]) with no tolerance, and the sub-row transforms take
    # their exact tie path for every row.
    fs = FS
    P = ocodes.samples_per_period(system, fs)
    L = 2 * P
    fc = grid(IF, 500.0, 500.0)
    systems = [(system, [3, 4], fc)]
    plan = tga.Plan(fs, 1, L, [tga.SystemSearch(n, f, p) for n, p, f in systems], input_format=1)
    packed = np.full((1, plan.row_bytes), 0xFF, dtype=np.uint8)
    peaks = plan.search(packed)
    levels = synth.unpack_levels(packed, L, 1)
    for pi, prn in enumerate(plan.prns[0]):
        code = ocodes.code_resampled(system, prn, float(P), L)
        ref, *_ = oracle_search(levels[0], fs, P, L, fc, code)
        assert_peak_matches(peaks[0][0, pi], ref, P, (False, False), what=f"{system} {prn}")
        assert int(peaks[0][0, pi]["best_freq_idx"]) == len(fc) - 1


def test_rows_matrix_flops_follow_the_plan(gpu, monkeypatch):
    """tga_plan_rows_matrix_flops counts the matrix-core instructions of the kernels a plan's
    groups run: 33 x 16 v_mfma_f32_16x16x4f32 per 16368-point (sub-)row with the matrix-core
    radix 31, none with the VALU codelet, 17 x 16 per 8184-point half sub-row, 16 per 496-point
    four-step row; the FP16 form 857 v_mfma_f32_16x16x32_f16 per (sub-)row."""
    f32_op, f16_op = 2.0 * 16 * 16 * 4, 2.0 * 16 * 16 * 32
    fc = grid(IF, 500.0, 500.0)
    rows = 1 * 3 * len(fc) * 2
    for v, per_row in ((0, 0), (1, 528)):
        monkeypatch.setenv("TGA_FUSED_VARIANT", str(v))
        plan = tga.Plan(FS, 1, 2 * 16368, [tga.SystemSearch("gps", fc, [1, 2, 3])])
        assert plan.rows_matrix_flops() == (rows * per_row * f32_op, 0.0)
    monkeypatch.delenv("TGA_FUSED_VARIANT")
    plan = tga.Plan(FS, 1, 2 * 16368, [tga.SystemSearch("gps", fc, [1, 2, 3])], precision="fp16")
    assert plan.rows_matrix_flops() == (0.0, rows * 857 * f16_op)
    long_rows = 1 * 2 * len(fc) * 2  # galileo, 2 PRNs, folded
    for v, per_row in (("2", 4 * 528), ("5", 8 * 272), ("3", 132 * 16), ("0", 0)):
        monkeypatch.setenv("TGA_SPLIT_VARIANT", v)
        plan = tga.Plan(FS, 1, 2 * 65472, [tga.SystemSearch("galileo", fc, [3, 11])])
        assert plan.rows_matrix_flops() == (long_rows * per_row * f32_op, 0.0), v
