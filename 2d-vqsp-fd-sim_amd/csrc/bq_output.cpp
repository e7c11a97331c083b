// bq_output.cpp - receiver files in the reference's byte format, written from
// chunks of device samples (host side, no device work).
//
// STORE_SEISMOGRAMS (store_seismograms.F90:47-154) and
// STORE_SEISMOGRAMS_PRESSURE (store_seismograms_pressure.F90:45-83) open,
// append one line to and close every receiver file at every step. Here one
// call writes a whole chunk of steps: each file is opened once per chunk and
// receives all of the chunk's lines, which are the same bytes in the same
// order. The receivers are independent, so they are shared out over host
// threads.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <algorithm>
#include <atomic>
#include <string>
#include <thread>
#include <vector>

#include "../../include/biotfd.h"

namespace {

// Fortran Ew.d edit (w = 13, d = 6) of a binary32 value, as the reference's
// runtime writes it: a sign or a blank, "0.", six correctly rounded digits,
// "E", a signed two-digit exponent, right-justified; NaN and infinities are
// spelled "NaN", "Inf", "-Inf", right-justified.
void edit_e13_6(float x, char *out /* 13 chars, no terminator */) {
  char body[32];
  const double v = (double)x;
  int n;
  if (std::isnan(v)) {
    n = std::snprintf(body, sizeof body, "NaN");
  } else if (std::isinf(v)) {
    n = std::snprintf(body, sizeof body, v < 0 ? "-Inf" : "Inf");
  } else {
    const bool neg = std::signbit(v);
    char digits[8];
    int exp10 = 0;
    if (v == 0.0) {
      std::memcpy(digits, "000000", 7);
    } else {
      char s[32];
      std::snprintf(s, sizeof s, "%.5e", std::fabs(v));  // d.ddddde[+-]xx, correctly rounded
      digits[0] = s[0];
      std::memcpy(digits + 1, s + 2, 5);
      digits[6] = 0;
      exp10 = std::atoi(s + 8) + 1;
    }
    n = std::snprintf(body, sizeof body, "%s0.%sE%c%02d", neg ? "-" : "", digits,
                      exp10 >= 0 ? '+' : '-', exp10 >= 0 ? exp10 : -exp10);
  }
  const int pad = 13 - n;
  std::memset(out, ' ', 13);
  std::memcpy(out + (pad > 0 ? pad : 0), body, (size_t)(n < 13 ? n : 13));
}

// One line "(k(1X,E13.6))" appended to s.
void put_line(std::string &s, const float *v, int k) {
  char f[13];
  for (int c = 0; c < k; ++c) {
    s.push_back(' ');
    edit_e13_6(v[c], f);
    s.append(f, 13);
  }
  s.push_back('\n');
}

bool append_file(const std::string &path, const std::string &data) {
  FILE *f = std::fopen(path.c_str(), "ab");
  if (!f) return false;
  const bool ok = std::fwrite(data.data(), 1, data.size(), f) == data.size();
  return std::fclose(f) == 0 && ok;
}

}  // namespace

extern "C" int bq_format_e13_6(float x, char *out) {
  if (!out) return -1;
  edit_e13_6(x, out);
  out[13] = 0;
  return 0;
}

extern "C" int bq_write_seismograms(const char *dir, const float *samples, int first_step,
                                    int nsteps, int mr, float dt, const int32_t *inside,
                                    int key_tlq, int key_tlp, float *state) {
  if (!dir || (!samples && nsteps > 0 && mr > 0) || !state || first_step < 1 || nsteps < 0 ||
      mr < 0)
    return -1;
  std::atomic<int> bad{0};
  auto one = [&](int j) {
    // SSU, SSW, SSQX, SSQZ of receiver j (MODULE WAVEFIELD): the previous
    // sample of each column, from which the time derivatives are formed
    float *prev = state + (size_t)j * 4;
    const bool in_u = !inside || inside[2 * (size_t)j] != 0;
    const bool in_w = !inside || inside[2 * (size_t)j + 1] != 0;
    std::string uw, q, st;
    uw.reserve((size_t)nsteps * 71);
    if (key_tlq) q.reserve((size_t)nsteps * 71);
    if (key_tlp) st.reserve((size_t)nsteps * 43);
    for (int s = 0; s < nsteps; ++s) {
      const float *o = samples + ((size_t)s * mr + j) * 6;  // U, -W, QX, -QZ, PRES, -TXX
      const float t = dt * (float)(first_step + s);          // DT*(T)
      // store_seismograms.F90:54-100: SPU = (S - SSU)/DT, SSU = S (likewise W,
      // QX, QZ) for nodes inside the grid; outside, the sample and its
      // derivative stay 0
      float d[4] = {0.0f, 0.0f, 0.0f, 0.0f};
      const bool in[4] = {in_u, in_w, in_u, in_w};
      for (int c = 0; c < 4; ++c) {
        if (!in[c]) continue;
        d[c] = (o[c] - prev[c]) / dt;
        prev[c] = o[c];
      }
      const float luw[5] = {t, prev[0], prev[1], d[0], d[1]};
      put_line(uw, luw, 5);
      if (key_tlq) {
        const float lq[5] = {t, prev[2], prev[3], d[2], d[3]};
        put_line(q, lq, 5);
      }
      if (key_tlp) {
        // store_seismograms_pressure.F90:50-56: PRES and -TXX at (IRECZ, LRECX)
        const float lp[3] = {t, o[4], o[5]};
        put_line(st, lp, 3);
      }
    }
    const std::string num = std::to_string(j + 1);
    const std::string base = std::string(dir) + "/";
    if (!append_file(base + "REC_UW_" + num + ".asc", uw)) bad.store(1);
    if (key_tlq && !append_file(base + "REC_Q_" + num + ".asc", q)) bad.store(1);
    if (key_tlp && !append_file(base + "REC_STRESS_" + num + ".asc", st)) bad.store(1);
  };
  // receivers in interleaved order over up to 16 threads
  const int hw = (int)std::thread::hardware_concurrency();
  const int nth = std::max(1, std::min(std::min(16, mr), hw));
  std::vector<std::thread> pool;
  for (int w = 1; w < nth; ++w)
    pool.emplace_back([&, w] { for (int j = w; j < mr; j += nth) one(j); });
  for (int j = 0; j < mr; j += nth) one(j);
  for (auto &t : pool) t.join();
  return bad.load() ? -10 : 0;
}
