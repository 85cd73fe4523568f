"""Index arithmetic the GPU kernels rely on, checked on the CPU with the formulas the kernels use
(search.hip): the prime-factor (Good-Thomas) order of the 16368-point rows, the lane sets of the
top-2 selection, and the tiled enumeration of the long-code rows.  These are properties of the
design, not of the reference; the GPU parity tests check the kernels themselves."""
import numpy as np

P = 16368  # 33 x 16 x 31
CW = 16    # ceil(16.368e6 / 1.023e6)


def pfa_nat(s):
    """natural index held at storage position s = 496 n1 + 31 n2 + n3"""
    n1, b = divmod(s, 496)
    n2, n3 = divmod(b, 31)
    return (496 * n1 + 1023 * n2 + 528 * n3) % P


def out_idx(j, r):
    """sample index of output r of pass-C butterfly j = k1 + 33 k2"""
    k2, k1 = divmod(j, 33)
    return (496 * k1 + 15345 * k2 + 528 * r) % P


def ring(a, b):
    d = abs(a - b) % P
    return min(d, P - d)


def test_prime_factor_passes_equal_the_dft():
    # the three passes of the rows kernels, with their LDS layouts and no twiddle factors
    rng = np.random.default_rng(3)
    x = rng.standard_normal(P) + 1j * rng.standard_normal(P)
    nat = np.array([pfa_nat(s) for s in range(P)])
    assert len(set(nat.tolist())) == P
    a = x[nat].reshape(33, 16, 31)                       # [n1][n2][n3]
    w = lambda r: np.exp(-2j * np.pi * np.outer(np.arange(r), np.arange(r)) / r)
    a = np.einsum("kn,nab->kab", w(33), a)              # pass A over n1 -> k1
    buf0 = np.zeros((16, 1023), complex)                # row n2, column 33 n3 + k1
    for k1 in range(33):
        buf0[:, 33 * np.arange(31) + k1] = a[k1]
    b = w(16) @ buf0                                     # pass B over n2 -> k2
    buf1 = np.zeros((31, 528), complex)                 # row n3, column k1 + 33 k2
    for c in range(1023):
        n3, k1 = divmod(c, 33)
        buf1[n3, k1 + 33 * np.arange(16)] = b[:, c]
    c = w(31) @ buf1                                     # pass C over n3 -> k3
    out = np.zeros(P, complex)
    for j in range(528):
        out[[out_idx(j, r) for r in range(31)]] = c[:, j]
    ref = np.fft.fft(x)
    assert np.abs(out - ref).max() <= 1e-9 * np.abs(ref).max()


def test_good_thomas_codelets():
    # Dft<33> = 3 x 11 and Dft<10> = 2 x 5 in registers: input (R2 n1 + R1 n2) mod R, output
    # (E1 k1 + E2 k2) mod R with E1 = 1 mod R1, 0 mod R2 and E2 = 0 mod R1, 1 mod R2
    rng = np.random.default_rng(4)
    for r1, r2 in ((3, 11), (2, 5)):
        r = r1 * r2
        e1 = next(e for e in range(r) if e % r1 == 1 and e % r2 == 0)
        e2 = next(e for e in range(r) if e % r1 == 0 and e % r2 == 1)
        x = rng.standard_normal(r) + 1j * rng.standard_normal(r)
        u = np.array([[x[(r2 * n1 + r1 * n2) % r] for n2 in range(r2)] for n1 in range(r1)])
        u = np.fft.fft(np.fft.fft(u, axis=1), axis=0)
        out = np.zeros(r, complex)
        for k1 in range(r1):
            for k2 in range(r2):
                out[(e1 * k1 + e2 * k2) % r] = u[k1, k2]
        assert np.allclose(out, np.fft.fft(x))


def test_top2_sets_are_spaced_beyond_the_window():
    # every top-2 set of every lane holds outputs pairwise more than 2 cw apart, so at most one
    # of them falls inside the window around the row's argmax
    def spacing(s):
        s = list(s)
        return min(ring(a, b) for i, a in enumerate(s) for b in s[i + 1:])
    worst = []
    # FP32, VALU radix-31: one butterfly per thread; left-over units (512 + b, kp / 31 - kp)
    worst += [spacing(out_idx(t, r) for r in range(31)) for t in range(512)]
    worst += [ring(out_idx(512 + b, kp), out_idx(512 + b, 31 - kp))
              for b in range(16) for kp in range(1, 16)]
    # FP32, matrix-core radix-31: tiles wave + 8 ts; sets {0, 1, 4} and {2, 3}
    for wave in range(8):
        for lane in range(64):
            ln16, gc = lane & 15, lane >> 4
            sets = {0: [], 31: []}
            for ts in range(5 if wave == 0 else 4):
                q = 16 * (wave + 8 * ts) + ln16
                for e in range(4):
                    kk = 4 * gc + e
                    slot = 31 if ts in (2, 3) else 0
                    sets[slot].append(out_idx(q, kk))
                    if kk:
                        sets[slot].append(out_idx(q, 31 - kk))
            worst += [spacing(v) for v in sets.values() if len(v) > 1]
    # FP16: butterflies j0 and j0 + 128 (j0 = 16 t0 + ln16, t0 = wave or wave + 8)
    for wave in range(16):
        for lane in range(64):
            ln16, grp = lane & 15, lane >> 4
            j0 = 16 * (wave if wave < 8 else wave + 8) + ln16
            outs = [out_idx(j, 8 * t + 2 * grp + e) for j in (j0, j0 + 128) for t in range(4)
                    for e in range(2) if t < 3 or grp < 3 or e == 0]
            worst.append(spacing(outs))
    assert min(worst) > 2 * CW


def big_item(i, n_prn, n_bin, n_ant, hs, tx=8, tp=8):
    """(xrow, prn) of enumerated long row i (k_big_pre, big_row)"""
    n_xe = n_ant * n_bin * 2 // hs
    blk = tx * n_prn
    xb, rem = divmod(i, blk)
    a = min(n_xe - xb * tx, tx)
    pt, rem = divmod(rem, a * tp)
    q, off = divmod(rem, a)
    xe = xb * tx + off
    n_ab = n_ant * n_bin
    xrow = ((xe - n_ab) * 2 + 1 if xe >= n_ab else xe * 2) if hs == 1 else xe * 2
    return xrow, pt * tp + q


def test_long_row_tiles_enumerate_every_row_once():
    for n_ant, n_bin, n_prn, hs in ((24, 201, 63, 1), (24, 81, 50, 1), (3, 5, 7, 2),
                                    (5, 7, 13, 1), (1, 1, 1, 2), (2, 3, 9, 1)):
        rows = {big_item(i, n_prn, n_bin, n_ant, hs)
                for i in range(n_ant * n_bin * 2 * n_prn // hs)}
        assert len(rows) == n_ant * n_bin * 2 * n_prn // hs
        assert all(0 <= x < n_ant * n_bin * 2 and 0 <= p < n_prn for x, p in rows)
        assert hs == 1 or all(x % 2 == 0 for x, _ in rows)


def _half_split_dft(z, R0):
    """The FP32 split's half sub-rows (k_half_pre + k_halfrows in search.hip) in numpy, with the
    kernels' index formulas: z (natural order, P = R0 x 16368) stored in the 16368-segment
    prime-factor order, the pre-pass gather of the (lo, lo + 248) pair per segment, the R-point
    DFT and the twiddles from the two table values, then the 8184-point transform's three passes
    with their LDS positions and output map.  Returns the P-point DFT of z."""
    S, R = 8184, 2 * R0
    Pl = R0 * P
    stored = np.empty(Pl, complex)
    s16 = np.arange(P)
    for g in range(R0):
        stored[g * P + s16] = z[g * P + np.array([pfa_nat(int(s)) for s in s16])]
    s8 = np.arange(S)
    c1, b = np.divmod(s8, 248)
    c2, c3 = np.divmod(b, 31)
    m = (496 * c1 + 7161 * c2 + 528 * c3) % S
    lo = 496 * c1 + 31 * ((8 - c2) & 7) + c3
    swap = ((-m) & 15) >= 8
    v = np.empty((R, S), complex)
    for g in range(R0):
        pl, ph = stored[g * P + lo], stored[g * P + lo + 248]
        v[2 * g] = np.where(swap, ph, pl)
        v[2 * g + 1] = np.where(swap, pl, ph)
    v = np.fft.fft(v, axis=0)  # R-point DFT over n0
    w1 = np.exp(-2j * np.pi * m / Pl)
    wh = np.exp(-2j * np.pi * (m * R0 % Pl) / Pl)
    wk = w1.copy()
    v[R0] *= wh
    for k0 in range(1, R0):
        v[k0] *= wk
        v[k0 + R0] *= wk * wh
        wk = wk * w1
    Z = np.empty(Pl, complex)
    ta = np.arange(248)
    j = np.arange(1023)
    q = np.arange(264)
    for k0 in range(R):
        y = v[k0]
        buf = np.empty(S, complex)
        a = np.fft.fft(y[ta[:, None] + 248 * np.arange(33)[None, :]], axis=1)  # pass A
        buf[(33 * ta)[:, None] + np.arange(33)[None, :]] = a
        bb = np.fft.fft(buf[j[:, None] + 1023 * np.arange(8)[None, :]], axis=1)  # pass B
        buf[j[:, None] + 1023 * np.arange(8)[None, :]] = bb
        base = q + 990 * (q // 33)  # c31_pos
        cc = np.fft.fft(buf[base[:, None] + 33 * np.arange(31)[None, :]], axis=1)  # pass C
        k2, k1 = np.divmod(q, 33)
        samp = (248 * k1[:, None] + 1023 * k2[:, None] + 264 * np.arange(31)[None, :]) % S
        Z[k0 + R * samp] = cc
    return Z


def test_half_subrows_equal_the_long_dft():
    # R0 = 4 (E1-C, 8 half sub-rows) and 10 (B1C / L1C, 20 half sub-rows)
    rng = np.random.default_rng(7)
    for R0 in (4, 10):
        z = rng.standard_normal(R0 * P) + 1j * rng.standard_normal(R0 * P)
        Z = _half_split_dft(z, R0)
        ref = np.fft.fft(z)
        assert np.max(np.abs(Z - ref)) < 1e-9 * np.max(np.abs(ref))


def test_half_subrow_maps_are_bijections():
    S = 8184
    s8 = np.arange(S)
    c1, b = np.divmod(s8, 248)
    c2, c3 = np.divmod(b, 31)
    m = (496 * c1 + 7161 * c2 + 528 * c3) % S
    assert np.array_equal(np.sort(m), s8)
    assert np.array_equal(m % 33, c1) and np.array_equal(m % 8, c2) and np.array_equal(m % 31, c3)
    q = np.arange(264)
    k2, k1 = np.divmod(q, 33)
    samp = (248 * k1[:, None] + 1023 * k2[:, None] + 264 * np.arange(31)[None, :]) % S
    assert np.array_equal(np.sort(samp.ravel()), s8)
    # the pre-pass's (lo, lo + 248) pairs cover every position of a 16368 segment once
    lo = 496 * c1 + 31 * ((8 - c2) & 7) + c3
    assert np.array_equal(np.sort(np.concatenate([lo, lo + 248])), np.arange(P))
