// Spreading-code generation for the six searched signals (host side; computed once per plan).
//
//   C/A Gold codes (GPS, SBAS, QZSS)  acquisition.rs:84-125, sbas.rs:39-44,106-118,
//                                     qzss.rs:40-44
//   Galileo E1-C memory codes         galileo.rs:82-105, galileo_codes.rs:20-70 (ICD data)
//   BeiDou B1Cp Weil codes            beidou_codes.rs:46-158
//   GPS L1Cp Weil codes               l1c_codes.rs:45-189
//   resampling                        acquisition.rs:146-159 (same rule in every module)
#include <cmath>
#include <mutex>

#include "internal.h"
#include "tables_gen.h"  // kGalileoE1cHex[50] (generated from tables/galileo_e1c_hex.txt)

namespace tga {

namespace {

// G2 delays in chips; the values are the published PRN assignments the reference uses
// (including its quirks: PRN 34 and PRN 37 both 950, PRN 38 delay 0).
const int kGpsDelay[38] = {5,   6,   7,   8,   17,  18,  139, 140, 141, 251, 252, 254, 255,
                           256, 257, 258, 469, 470, 471, 472, 473, 474, 509, 512, 513, 514,
                           515, 516, 859, 860, 861, 862, 863, 950, 947, 948, 950, 0};
const int kSbasDelay[39] = {145, 175, 52,  21,  237, 235, 886,  657, 634, 762, 355, 1012, 176,
                            603, 130, 359, 595, 68,  386, 797,  456, 499, 883, 307, 127,  211,
                            121, 118, 163, 628, 853, 484, 289,  811, 202, 1021, 463, 568, 904};
const int kQzssDelay[23] = {476, 193, 109, 445, 291, 87,  399, 292, 901, 339, 208, 711,
                            189, 263, 537, 663, 942, 173, 900, 30,  500, 935, 556};

// Weil code parameters (w, p), PRN 1..63.
const int kB1cp[63][2] = {
    {796, 7575},  {156, 2369},  {4198, 5688}, {3941, 539},  {1374, 2270}, {1338, 7306},
    {1833, 6457}, {2521, 6254}, {3175, 5644}, {168, 7119},  {2715, 1402}, {4408, 5557},
    {3160, 5764}, {2796, 1073}, {459, 7001},  {3594, 5910}, {4813, 10060}, {586, 2710},
    {1428, 1546}, {2371, 6887}, {2285, 1883}, {3377, 5613}, {4965, 5062}, {3779, 1038},
    {4547, 10170}, {1646, 6484}, {1430, 1718}, {607, 2535}, {2118, 1158}, {4709, 526},
    {1149, 7331}, {3283, 5844}, {2473, 6423}, {1006, 6968}, {3670, 1280}, {1817, 1838},
    {771, 1989},  {2173, 6468}, {740, 2091},  {1433, 1581}, {2458, 1453}, {3459, 6252},
    {2155, 7122}, {1205, 7711}, {413, 7216},  {874, 2113},  {2463, 1095}, {1106, 1628},
    {1590, 1713}, {3873, 6102}, {4026, 6123}, {4272, 6070}, {3556, 1115}, {128, 8047},
    {1200, 6795}, {130, 2575},  {4494, 53},   {1871, 1729}, {3073, 6388}, {4386, 682},
    {4098, 5565}, {1923, 7160}, {1176, 2277}};
const int kL1cp[63][2] = {
    {5111, 412},  {5109, 161},  {5108, 1},    {5106, 303},  {5103, 207},  {5101, 4971},
    {5100, 4496}, {5098, 5},    {5095, 4557}, {5094, 485},  {5093, 253},  {5091, 4676},
    {5090, 1},    {5081, 66},   {5080, 4485}, {5069, 282},  {5068, 193},  {5054, 5211},
    {5044, 729},  {5027, 4848}, {5026, 982},  {5014, 5955}, {5004, 9805}, {4980, 670},
    {4915, 464},  {4909, 29},   {4893, 429},  {4885, 394},  {4832, 616},  {4824, 9457},
    {4591, 4429}, {3706, 4771}, {5092, 365},  {4986, 9705}, {4965, 9489}, {4920, 4193},
    {4917, 9947}, {4858, 824},  {4847, 864},  {4790, 347},  {4770, 677},  {4318, 6544},
    {4126, 6312}, {3961, 9804}, {3790, 278},  {4911, 9461}, {4881, 444},  {4827, 4839},
    {4795, 4144}, {4789, 9875}, {4725, 197},  {4675, 1156}, {4539, 4674}, {4535, 10035},
    {4458, 4504}, {4197, 5},    {4096, 9937}, {3484, 430},  {3481, 5},    {3393, 355},
    {3175, 909},  {2360, 1622}, {1852, 6284}};

const SystemInfo kSystems[TGA_NUM_SYSTEMS] = {
    {TGA_SYS_GPS, "gps", "GPS%02d", 1, 38, 0.001, 1023, TGA_ACR_GPS_L1_CA},
    {TGA_SYS_GALILEO, "galileo", "GSAT%02d", 1, 50, 0.004, 4092, TGA_ACR_GALILEO_E1C},
    {TGA_SYS_BEIDOU, "beidou", "BEIDOU%02d", 1, 63, 0.010, 20460, TGA_ACR_BEIDOU_B1C},
    {TGA_SYS_SBAS, "sbas", "SBAS%03d", 120, 158, 0.001, 1023, TGA_ACR_GPS_L1_CA},
    {TGA_SYS_L1C, "l1c", "GPSL1C%02d", 1, 63, 0.010, 20460, TGA_ACR_GPS_L1C},
    {TGA_SYS_QZSS, "qzss", "QZSS%03d", 184, 206, 0.001, 1023, TGA_ACR_GPS_L1_CA},
};

// Ten-stage LFSR sequence of one C/A period: all-ones start, output from stage 10,
// feedback = XOR of the tapped stages.  Bits packed in a uint16 state, stage s in bit s-1.
std::vector<uint8_t> lfsr_1023(uint16_t tap_mask) {
  std::vector<uint8_t> out(1023);
  uint16_t st = 0x3FF;
  for (int i = 0; i < 1023; ++i) {
    out[i] = (st >> 9) & 1;
    uint16_t fb = __builtin_parity(st & tap_mask);
    st = static_cast<uint16_t>(((st << 1) | fb) & 0x3FF);
  }
  return out;
}

const std::vector<uint8_t>& g1_seq() {
  static const std::vector<uint8_t> g = lfsr_1023((1u << 2) | (1u << 9));  // stages 3, 10
  return g;
}
const std::vector<uint8_t>& g2_seq() {  // stages 2, 3, 6, 8, 9, 10
  static const std::vector<uint8_t> g =
      lfsr_1023((1u << 1) | (1u << 2) | (1u << 5) | (1u << 7) | (1u << 8) | (1u << 9));
  return g;
}

// The reference forms -(g1*g2) on bipolar (+1 <-> bit 0) values, i.e. +1 where g1 XOR g2 = 1,
// with G2 delayed by `shift` chips.
std::vector<int8_t> ca_from_delay(int shift) {
  const auto& g1 = g1_seq();
  const auto& g2 = g2_seq();
  std::vector<int8_t> out(1023);
  for (int i = 0; i < 1023; ++i) {
    int j = ((i - shift) % 1023 + 1023) % 1023;
    out[i] = (g1[i] ^ g2[j]) ? 1 : -1;
  }
  return out;
}

std::vector<uint8_t> legendre(int p) {
  std::vector<uint8_t> l(p, 0);
  for (int64_t x = 1; x < p; ++x) l[(x * x) % p] = 1;
  l[0] = 0;
  return l;
}

const std::vector<uint8_t>& legendre_b1c() {
  static const std::vector<uint8_t> l = legendre(10243);
  return l;
}
const std::vector<uint8_t>& legendre_l1c() {
  static const std::vector<uint8_t> l = legendre(10223);
  return l;
}

// [+c, -c] half-chips with 0 -> +1, 1 -> -1.
std::vector<int8_t> boc_from_bits(const std::vector<uint8_t>& bits) {
  std::vector<int8_t> out(2 * bits.size());
  for (size_t i = 0; i < bits.size(); ++i) {
    int8_t c = bits[i] ? -1 : 1;
    out[2 * i] = c;
    out[2 * i + 1] = static_cast<int8_t>(-c);
  }
  return out;
}

std::vector<int8_t> b1c_pilot(int prn) {
  const int n = 10243;
  const auto& L = legendre_b1c();
  const int w = kB1cp[prn - 1][0], p = kB1cp[prn - 1][1];
  std::vector<uint8_t> bits(10230);
  for (int i = 0; i < 10230; ++i) {
    int k = (i + p - 1) % n;
    bits[i] = L[k] ^ L[(k + w) % n];
  }
  return boc_from_bits(bits);
}

std::vector<int8_t> l1c_pilot(int prn) {
  const int n = 10223;
  const auto& L = legendre_l1c();
  const int w = kL1cp[prn - 1][0], p = kL1cp[prn - 1][1];
  static const uint8_t kExpansion[7] = {0, 1, 1, 0, 1, 0, 0};
  std::vector<uint8_t> bits;
  bits.reserve(10230);
  for (int k = 0; k < p - 1; ++k) bits.push_back(L[k] ^ L[(k + w) % n]);
  for (uint8_t b : kExpansion) bits.push_back(b);
  for (int k = p - 1; k < n; ++k) bits.push_back(L[k] ^ L[(k + w) % n]);
  return boc_from_bits(bits);
}

std::vector<int8_t> e1c(int prn) {
  const char* hex = kGalileoE1cHex[prn - 1];
  std::vector<int8_t> out(4092);
  for (int i = 0; i < 1023; ++i) {
    char ch = hex[i];
    int v = (ch >= '0' && ch <= '9') ? ch - '0' : (ch >= 'A' && ch <= 'F') ? ch - 'A' + 10
                                                                              : ch - 'a' + 10;
    for (int b = 0; b < 4; ++b) out[4 * i + b] = ((v >> (3 - b)) & 1) ? 1 : -1;
  }
  return out;
}

}  // namespace

const SystemInfo& system_info(int system) {
  if (system < 0 || system >= TGA_NUM_SYSTEMS)
    fail(TGA_ERR_INVALID_ARGUMENT, "unknown system id " + std::to_string(system));
  return kSystems[system];
}

void check_prn(int system, int prn) {
  const SystemInfo& s = system_info(system);
  if (prn < s.prn_first || prn > s.prn_last) {
    char buf[160];
    std::snprintf(buf, sizeof buf, "%s PRN must be in %d..=%d, got %d", s.key, s.prn_first,
                  s.prn_last, prn);
    fail(TGA_ERR_PRN_RANGE, buf);
  }
}

std::vector<int8_t> code_chips(int system, int prn) {
  check_prn(system, prn);
  switch (system) {
    case TGA_SYS_GPS: return ca_from_delay(kGpsDelay[prn - 1]);
    case TGA_SYS_SBAS: return ca_from_delay(kSbasDelay[prn - 120]);
    case TGA_SYS_QZSS: return ca_from_delay(kQzssDelay[prn - 184]);
    case TGA_SYS_GALILEO: return e1c(prn);
    case TGA_SYS_BEIDOU: return b1c_pilot(prn);
    case TGA_SYS_L1C: return l1c_pilot(prn);
  }
  fail(TGA_ERR_INVALID_ARGUMENT, "unknown system");
}

std::vector<int8_t> replica_chips(int system, int prn, int replica) {
  if (replica == TGA_REPLICA_REFERENCE) return code_chips(system, prn);
  if (replica != TGA_REPLICA_E1C_BOC11)
    fail(TGA_ERR_INVALID_ARGUMENT, "unknown replica " + std::to_string(replica));
  if (system != TGA_SYS_GALILEO)
    fail(TGA_ERR_UNSUPPORTED, "the BOC(1,1) replica is defined for Galileo E1-C only");
  const std::vector<int8_t> c = code_chips(system, prn);
  std::vector<int8_t> out(2 * c.size());
  for (size_t i = 0; i < c.size(); ++i) {  // sine-phased subcarrier: + then - within a chip
    out[2 * i] = c[i];
    out[2 * i + 1] = static_cast<int8_t>(-c[i]);
  }
  return out;
}

void resample(const std::vector<int8_t>& chips, double samples_per_code, size_t n,
              double* out) {
  const size_t L = chips.size();
  const double spchip = samples_per_code / static_cast<double>(L);
  for (size_t i = 0; i < n; ++i) {
    double q = std::floor(static_cast<double>(i) / spchip);
    size_t idx = static_cast<size_t>(q) % L;
    out[i] = chips[idx];
  }
}

size_t samples_per_period(int system, double fs) {
  const SystemInfo& s = system_info(system);
  if (s.period_s == 0.001) return static_cast<size_t>(fs / 1000.0);  // acquisition.rs:272-273
  return static_cast<size_t>(fs * s.period_s);                       // galileo.rs:230 etc.
}

}  // namespace tga
