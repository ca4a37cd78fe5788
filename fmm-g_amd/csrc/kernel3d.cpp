// Host dense kernel-matrix builder (kernel3d_mpi.cpp:142-177 buildKnlIntCtx).
#include "kernel3d.hpp"

#include <cstdint>

namespace kifmm {

template <int KT>
static void build_typed(const KParams<double>& p, int ns, const double* spos, const double* snor,
                        int nt, const double* tpos, double* mat) {
  constexpr int SD = src_dof(KT), TD = trg_dof(KT);
  const int64_t M = int64_t(nt) * TD;
  double blk[SD * TD > 0 ? SD * TD : 1];
  for (int j = 0; j < ns; j++)
    for (int i = 0; i < nt; i++) {
      kblock<KT, double>(p, tpos[3 * i] - spos[3 * j], tpos[3 * i + 1] - spos[3 * j + 1],
                         tpos[3 * i + 2] - spos[3 * j + 2], snor[3 * j], snor[3 * j + 1],
                         snor[3 * j + 2], blk);
      for (int t = 0; t < TD; t++)
        for (int s = 0; s < SD; s++) mat[int64_t(i * TD + t) + int64_t(j * SD + s) * M] = blk[t * SD + s];
    }
}

int build_kernel_matrix(int kt, const double* coefs, int ns, const double* spos,
                        const double* snor, int nt, const double* tpos, double* mat) {
  KParams<double> p = make_params(kt, coefs);
  switch (kt) {
#define KIFMM_CASE(K) case K: build_typed<K>(p, ns, spos, snor, nt, tpos, mat); return 0;
    KIFMM_FOR_EACH_KERNEL(KIFMM_CASE)
#undef KIFMM_CASE
    default: return -1;
  }
}

}  // namespace kifmm
