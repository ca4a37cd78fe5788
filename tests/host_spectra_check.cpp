// Host checks of the mixed path's spectrum bookkeeping, built and run by
// tests/test_host_spectra.py against fmm-g_amd/lib/libkifmm.so:
//  * spec_pack maps the r2c half spectrum onto [0, spec_packed_count) one to one
//    for the stored values, every other value to the conjugate of a stored one
//    at its mirror (-ka, -kb) in the planes kc = 0, np;
//  * SpecLine (the per-line form the transforms use) agrees with spec_pack;
//  * pack_spectra32 puts each stored value at its packed index and zeros in the
//    row padding;
//  * the operator tables: table_sym3 holds for the Stokes and Navier equivalent
//    kernels and fails once one entry is perturbed; table_mirror fails on the
//    tables with the wrap displacement and holds on the wrap-free tables of the
//    even kernels (Laplace, Navier), not on Stokes (odd fourth column).
// Prints one line per failed check and "OK" when none failed.
#include <cmath>
#include <cstdio>
#include <vector>

#include "../fmm-g_amd/csrc/device.hpp"
#include "../fmm-g_amd/csrc/fmm3d.hpp"
#include "../fmm-g_amd/csrc/matmgnt3d.hpp"

using namespace kifmm;

static int fails = 0;
#define CHECK(cond, ...)          \
  do {                            \
    if (!(cond)) {                \
      fails++;                    \
      std::printf(__VA_ARGS__);   \
      std::printf("\n");          \
    }                             \
  } while (0)

template <int NP>
static void check_layout() {
  const int n = 2 * NP, hh = NP + 1, cnt = spec_packed_count(NP), len = spec_packed_len(NP);
  CHECK(len % 16 == 0 && len >= cnt && len - cnt < 16, "np %d: packed length %d for %d values", NP, len, cnt);
  std::vector<int> owner(cnt, -1);
  for (int ka = 0; ka < n; ka++)
    for (int kb = 0; kb < n; kb++)
      for (int kc = 0; kc < hh; kc++) {
        const int code = spec_pack(NP, ka, kb, kc), idx = code >> 1, cj = code & 1;
        CHECK(idx >= 0 && idx < cnt, "np %d: index %d out of range at (%d,%d,%d)", NP, idx, ka, kb, kc);
        if (idx < 0 || idx >= cnt) continue;
        if (!cj) {
          CHECK(owner[idx] < 0, "np %d: index %d stored twice", NP, idx);
          owner[idx] = (ka * n + kb) * hh + kc;
        } else {
          CHECK(kc == 0 || kc == NP, "np %d: conjugate outside the mirrored planes", NP);
          CHECK(spec_pack(NP, (n - ka) % n, (n - kb) % n, kc) == (idx << 1), "np %d: mirror of (%d,%d,%d) not stored",
                NP, ka, kb, kc);
        }
        const SpecLine<NP> L(kb * hh + kc);
        CHECK(L.index(ka) == idx && int(L.conj(ka)) == cj, "np %d: SpecLine differs at (%d,%d,%d)", NP, ka, kb, kc);
      }
  for (int i = 0; i < cnt; i++) CHECK(owner[i] >= 0, "np %d: index %d never stored", NP, i);
  // pack_spectra32: two rows of distinct values
  const int h = n * n * hh;
  std::vector<double> full(size_t(2) * h * 2);
  for (int r = 0; r < 2; r++)
    for (int f = 0; f < h; f++) {
      full[(size_t(r) * h + f) * 2] = 1.0 + r * h + f;
      full[(size_t(r) * h + f) * 2 + 1] = -(1.0 + f);
    }
  const std::vector<float> pk = pack_spectra32(full, NP);
  CHECK(pk.size() == size_t(2) * len * 2, "np %d: packed table size %zu", NP, pk.size());
  for (int r = 0; r < 2 && pk.size() == size_t(2) * len * 2; r++) {
    for (int i = 0; i < cnt; i++) {
      const int f = owner[i];
      CHECK(pk[(size_t(r) * len + i) * 2] == float(1.0 + r * h + f) && pk[(size_t(r) * len + i) * 2 + 1] == float(-(1.0 + f)),
            "np %d: packed value %d of row %d", NP, i, r);
    }
    for (int i = cnt; i < len; i++)
      CHECK(pk[(size_t(r) * len + i) * 2] == 0.f && pk[(size_t(r) * len + i) * 2 + 1] == 0.f, "np %d: padding %d", NP, i);
  }
}

static void check_tables(int kt, std::vector<double> cf, int np, bool even, bool sym) {
  std::string err;
  auto op = get_operators(kt, cf, np, &err);
  CHECK(op != nullptr, "kernel %d np %d: %s", kt, np, err.c_str());
  if (!op) return;
  const int sdtd = op->sd * op->td;
  CHECK(!table_mirror(op->m2l, sdtd, op->H), "kernel %d np %d: table with the wrap displacement is mirror-symmetric", kt, np);
  const std::vector<double> nw = m2l_table_nowrap(*op);
  CHECK(nw.size() == op->m2l.size(), "kernel %d np %d: wrap-free table size", kt, np);
  CHECK(table_mirror(nw, sdtd, op->H) == even, "kernel %d np %d: wrap-free mirror check gives %d", kt, np, int(!even));
  CHECK(table_sym3(op->m2l, op->sd, op->td, op->H) == sym, "kernel %d np %d: sym3 check gives %d", kt, np, int(!sym));
  if (sym) {
    std::vector<double> t = op->m2l;
    // DOF pair (1, 0) of slot (3, 3, 3), frequency 5, real part
    const size_t at = ((size_t(Operators::slot(3, 3, 3)) * sdtd + size_t(1 * op->td + 0)) * op->H + 5) * 2;
    t[at] += 1e-6 * (std::fabs(t[at]) + 1.0);
    CHECK(!table_sym3(t, op->sd, op->td, op->H), "kernel %d np %d: sym3 check misses a perturbed entry", kt, np);
  }
}

int main() {
  check_layout<3>();
  check_layout<4>();
  check_layout<5>();
  check_layout<6>();
  check_layout<7>();
  check_layout<8>();
  check_tables(111, {1.0, 0.25}, 4, true, false);
  check_tables(111, {1.0, 0.25}, 6, true, false);
  check_tables(301, {1.0, 0.25}, 4, false, true);
  check_tables(511, {1.0, 0.3}, 4, true, true);
  if (!fails) std::printf("OK\n");
  return fails ? 1 : 0;
}
