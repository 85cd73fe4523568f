"""Spreading-code generators and replica resampling (oracle restatement).

Follows the reference's generators:
  GPS/SBAS/QZSS C/A Gold codes   acquisition.rs:84-125, sbas.rs:39-44, qzss.rs:40-44
  Galileo E1-C memory codes      galileo.rs:82-105 (+ galileo_codes.rs:20-70 data)
  BeiDou B1C pilot Weil codes    beidou_codes.rs:46-158
  GPS L1C pilot Weil codes       l1c_codes.rs:45-189
  Replica resampling             acquisition.rs:146-159, galileo.rs:112-122,
                                 beidou_codes.rs:166-177, l1c_codes.rs:197-208
Chips are returned as int8 arrays of +1/-1 in the reference's sign convention.
"""
import os

import numpy as np

_TABLE_DIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                          "tart-gnss_amd", "csrc", "tables")

# G2 delays (chips) -- acquisition.rs:84-87 (PRN 1..38, PRN 34 and 37 share 950, PRN 38 is 0)
GPS_G2_DELAY = [5, 6, 7, 8, 17, 18, 139, 140, 141, 251, 252, 254, 255, 256, 257, 258,
                469, 470, 471, 472, 473, 474, 509, 512, 513, 514, 515, 516, 859, 860,
                861, 862, 863, 950, 947, 948, 950, 0]
# sbas.rs:39-44 (PRN 120..158)
SBAS_G2_DELAY = [145, 175, 52, 21, 237, 235, 886, 657, 634, 762, 355, 1012, 176, 603, 130,
                 359, 595, 68, 386, 797, 456, 499, 883, 307, 127, 211, 121, 118, 163, 628,
                 853, 484, 289, 811, 202, 1021, 463, 568, 904]
# qzss.rs:40-44 (PRN 184..206)
QZSS_G2_DELAY = [476, 193, 109, 445, 291, 87, 399, 292, 901, 339, 208, 711, 189, 263, 537,
                 663, 942, 173, 900, 30, 500, 935, 556]

# (w, p) Weil parameters -- beidou_codes.rs:46-63
B1CP_PARAMS = [(796, 7575), (156, 2369), (4198, 5688), (3941, 539), (1374, 2270),
               (1338, 7306), (1833, 6457), (2521, 6254), (3175, 5644), (168, 7119),
               (2715, 1402), (4408, 5557), (3160, 5764), (2796, 1073), (459, 7001),
               (3594, 5910), (4813, 10060), (586, 2710), (1428, 1546), (2371, 6887),
               (2285, 1883), (3377, 5613), (4965, 5062), (3779, 1038), (4547, 10170),
               (1646, 6484), (1430, 1718), (607, 2535), (2118, 1158), (4709, 526),
               (1149, 7331), (3283, 5844), (2473, 6423), (1006, 6968), (3670, 1280),
               (1817, 1838), (771, 1989), (2173, 6468), (740, 2091), (1433, 1581),
               (2458, 1453), (3459, 6252), (2155, 7122), (1205, 7711), (413, 7216),
               (874, 2113), (2463, 1095), (1106, 1628), (1590, 1713), (3873, 6102),
               (4026, 6123), (4272, 6070), (3556, 1115), (128, 8047), (1200, 6795),
               (130, 2575), (4494, 53), (1871, 1729), (3073, 6388), (4386, 682),
               (4098, 5565), (1923, 7160), (1176, 2277)]
# (w, p) Weil parameters -- l1c_codes.rs:45-62
L1CP_PARAMS = [(5111, 412), (5109, 161), (5108, 1), (5106, 303), (5103, 207), (5101, 4971),
               (5100, 4496), (5098, 5), (5095, 4557), (5094, 485), (5093, 253), (5091, 4676),
               (5090, 1), (5081, 66), (5080, 4485), (5069, 282), (5068, 193), (5054, 5211),
               (5044, 729), (5027, 4848), (5026, 982), (5014, 5955), (5004, 9805), (4980, 670),
               (4915, 464), (4909, 29), (4893, 429), (4885, 394), (4832, 616), (4824, 9457),
               (4591, 4429), (3706, 4771), (5092, 365), (4986, 9705), (4965, 9489),
               (4920, 4193), (4917, 9947), (4858, 824), (4847, 864), (4790, 347), (4770, 677),
               (4318, 6544), (4126, 6312), (3961, 9804), (3790, 278), (4911, 9461),
               (4881, 444), (4827, 4839), (4795, 4144), (4789, 9875), (4725, 197),
               (4675, 1156), (4539, 4674), (4535, 10035), (4458, 4504), (4197, 5),
               (4096, 9937), (3484, 430), (3481, 5), (3393, 355), (3175, 909), (2360, 1622),
               (1852, 6284)]
L1C_EXPANSION = [0, 1, 1, 0, 1, 0, 0]  # l1c_codes.rs:65

CA_CHIPS = 1023
E1C_CHIPS = 4092
WEIL_CHIPS = 10230


def _lfsr_bits(taps, n=CA_CHIPS):
    """Binary 10-stage LFSR, all-ones start, output stage 10, feedback = XOR of ``taps``
    (1-based stage numbers).  Equivalent to the reference's bipolar product form."""
    reg = [1] * 10
    out = np.empty(n, dtype=np.int8)
    for i in range(n):
        out[i] = reg[9]
        fb = 0
        for t in taps:
            fb ^= reg[t - 1]
        reg = [fb] + reg[:9]
    return out


_G1 = _lfsr_bits((3, 10))
_G2 = _lfsr_bits((2, 3, 6, 8, 9, 10))


def ca_chips_from_delay(g2shift):
    """acquisition.rs:96-125: chip = -(g1*g2) in bipolar form, i.e. +1 where g1 XOR g2 = 1."""
    g2 = np.roll(_G2, int(g2shift))
    x = _G1 ^ g2
    return (2 * x.astype(np.int16) - 1).astype(np.int8)


def gps_ca(prn):
    if not 1 <= prn <= 38:
        raise ValueError(f"PRN must be in 1..=38, got {prn}")  # acquisition.rs:131-134
    return ca_chips_from_delay(GPS_G2_DELAY[prn - 1])


def sbas_ca(prn):
    if not 120 <= prn <= 158:
        raise ValueError(f"SBAS PRN must be in 120..=158, got {prn}")  # sbas.rs:112-116
    return ca_chips_from_delay(SBAS_G2_DELAY[prn - 120])


def qzss_ca(prn):
    if not 184 <= prn <= 206:
        raise ValueError(f"QZSS PRN must be in 184..=206, got {prn}")  # qzss.rs:112-116
    return ca_chips_from_delay(QZSS_G2_DELAY[prn - 184])


def _load_e1c_hex():
    rows = []
    with open(os.path.join(_TABLE_DIR, "galileo_e1c_hex.txt")) as f:
        for line in f:
            line = line.strip()
            if line and not line.startswith("#"):
                rows.append(line)
    return rows


_E1C_HEX = None


def e1c(prn):
    """galileo.rs:82-105: hex digits MSB-first, bit 1 -> +1, bit 0 -> -1."""
    global _E1C_HEX
    if not 1 <= prn <= 50:
        raise ValueError(f"Galileo PRN must be in 1..=50, got {prn}")
    if _E1C_HEX is None:
        _E1C_HEX = _load_e1c_hex()
    digits = np.array([int(c, 16) for c in _E1C_HEX[prn - 1]], dtype=np.int16)
    bits = ((digits[:, None] >> np.array([3, 2, 1, 0])) & 1).reshape(-1)[:E1C_CHIPS]
    return (2 * bits - 1).astype(np.int8)


def legendre(p):
    """L[0] = 0, L[i] = 1 iff i is a nonzero quadratic residue mod prime p."""
    seq = np.zeros(p, dtype=np.int8)
    x = np.arange(1, p, dtype=np.int64)
    seq[(x * x) % p] = 1
    seq[0] = 0
    return seq


def b1cp_bits(prn):
    """beidou_codes.rs:113-126: c[n] = W[(n + p - 1) mod 10243], W[k] = L[k] ^ L[(k + w) mod N]."""
    if not 1 <= prn <= 63:
        raise ValueError(f"BeiDou PRN must be in 1..=63, got {prn}")
    n_leg = 10243
    w, p = B1CP_PARAMS[prn - 1]
    leg = legendre(n_leg)
    k = (np.arange(WEIL_CHIPS) + p - 1) % n_leg
    return leg[k] ^ leg[(k + w) % n_leg]


def l1cp_bits(prn):
    """l1c_codes.rs:118-157: Weil code of length 10223 with 0110100 inserted before index p-1."""
    if not 1 <= prn <= 63:
        raise ValueError(f"GPS L1C PRN must be in 1..=63, got {prn}")
    n_leg = 10223
    w, p = L1CP_PARAMS[prn - 1]
    leg = legendre(n_leg)
    k = np.arange(n_leg)
    weil = leg[k] ^ leg[(k + w) % n_leg]
    return np.concatenate([weil[:p - 1], np.array(L1C_EXPANSION, np.int8), weil[p - 1:]])


def bits_to_bipolar_weil(bits):
    """0 -> +1, 1 -> -1 (beidou_codes.rs:131-138, l1c_codes.rs:162-169)."""
    return (1 - 2 * bits.astype(np.int16)).astype(np.int8)


def boc11(chips):
    """[+c, -c] per chip (beidou_codes.rs:145-158, l1c_codes.rs:175-189)."""
    out = np.empty(2 * len(chips), dtype=np.int8)
    out[0::2] = chips
    out[1::2] = -chips
    return out


def b1c_pilot(prn):
    return boc11(bits_to_bipolar_weil(b1cp_bits(prn)))


def l1c_pilot(prn):
    return boc11(bits_to_bipolar_weil(l1cp_bits(prn)))


SYSTEMS = {
    # name: (chip generator, code period seconds, label format, ACR table)
    "gps": (gps_ca, 0.001, "GPS{:02d}", "GPS_L1_CA_ACR_TABLE"),
    "sbas": (sbas_ca, 0.001, "SBAS{:03d}", "GPS_L1_CA_ACR_TABLE"),
    "qzss": (qzss_ca, 0.001, "QZSS{:03d}", "GPS_L1_CA_ACR_TABLE"),
    "galileo": (e1c, 0.004, "GSAT{:02d}", "GALILEO_E1C_ACR_TABLE"),
    "beidou": (b1c_pilot, 0.010, "BEIDOU{:02d}", "BEIDOU_B1C_ACR_TABLE"),
    "l1c": (l1c_pilot, 0.010, "GPSL1C{:02d}", "GPS_L1C_ACR_TABLE"),
}

PRN_RANGES = {
    "gps": range(1, 39), "sbas": range(120, 159), "qzss": range(184, 207),
    "galileo": range(1, 51), "beidou": range(1, 64), "l1c": range(1, 64),
}


def chips(system, prn):
    return SYSTEMS[system][0](prn)


def e1c_boc11(prn):
    """The BOC(1,1) E1-C replica of the plans' replica option (an extension: the reference
    correlates E1-C with plain chips, galileo.rs:112-122): each E1-C chip as [+c, -c]."""
    return boc11(e1c(prn))


def resample(chip_seq, samples_per_code, num_samples):
    """acquisition.rs:146-159: c[n] = chips[floor(n / (spc / L)) mod L] in f64 arithmetic."""
    L = len(chip_seq)
    spchip = float(samples_per_code) / float(L)
    idx = np.floor(np.arange(num_samples, dtype=np.float64) / spchip).astype(np.int64) % L
    return chip_seq[idx].astype(np.float64)


def code_resampled(system, prn, samples_per_code, num_samples):
    return resample(chips(system, prn), samples_per_code, num_samples)


def samples_per_period(system, fs):
    """(fs * period) as usize -- acquisition.rs:272-273, galileo.rs:230, beidou.rs:182."""
    if system in ("gps", "sbas", "qzss"):
        return int(fs / 1000.0)
    return int(fs * SYSTEMS[system][1])
