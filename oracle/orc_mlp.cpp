// ORACLE - test infrastructure only (see orc_math.h).
#include "orc_mlp.h"

#include <cmath>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

template <class R>
void mlp_encode(const float* rec, R* f) {
  const R pi = (R)3.14159265358979323846;
  int k = 0;
  for (int c = 0; c < 3; ++c) {
    const R p = (R)rec[c];
    for (int l = 0; l < 6; ++l) {
      const R a = (R)(1 << l) * pi * p;
      f[k++] = std::sin(a);
      f[k++] = std::cos(a);
    }
  }
  for (int v = 0; v < 2; ++v) {
    const R x = (R)rec[3 + 3 * v], y = (R)rec[4 + 3 * v], z = (R)rec[5 + 3 * v];
    const R zc = std::fmin((R)1, std::fmax((R)-1, z));
    f[k++] = std::acos(zc) / pi;
    f[k++] = std::atan2(y, x) / ((R)2 * pi) + (R)0.5;
  }
  f[k++] = (R)rec[9];
  f[k++] = (R)rec[10];
  f[k++] = (R)rec[11];
  f[k++] = (R)1;
  while (k < kMlpFeat) f[k++] = 0;
}

template <class R>
static void forward_one(const float* w, const float* rec, R h[kMlpHidden + 1][kMlpFeat], R z_out[kMlpOut]) {
  mlp_encode<R>(rec, h[0]);
  for (int l = 0; l < kMlpHidden; ++l) {
    const float* W = w + (size_t)l * kMlpFeat * kMlpFeat;
    for (int o = 0; o < kMlpFeat; ++o) {
      R acc = 0;
      for (int i = 0; i < kMlpFeat; ++i) acc += (R)W[o * kMlpFeat + i] * h[l][i];
      h[l + 1][o] = acc > 0 ? acc : (R)0;
    }
  }
  const float* Wo = w + (size_t)kMlpHidden * kMlpFeat * kMlpFeat;
  for (int o = 0; o < kMlpOut; ++o) {
    R acc = 0;
    for (int i = 0; i < kMlpFeat; ++i) acc += (R)Wo[o * kMlpFeat + i] * h[kMlpHidden][i];
    z_out[o] = acc;
  }
}

template <class R>
void mlp_forward(const float* w, const float* in, int n, R* out) {
#pragma omp parallel for schedule(static)
  for (int s = 0; s < n; ++s) {
    R h[kMlpHidden + 1][kMlpFeat], z[kMlpOut];
    forward_one<R>(w, in + (size_t)s * kMlpIn, h, z);
    for (int o = 0; o < kMlpOut; ++o) out[(size_t)s * kMlpOut + o] = std::exp(z[o]);
  }
}

template <class R>
double mlp_loss_grad(const float* w, const float* in, const float* target, int n, int loss, double exp_floor,
                     R* grad) {
  std::memset(grad, 0, sizeof(R) * kMlpWeights);
  double total = 0.0;
  const double inv_n = 1.0 / ((double)n * kMlpOut);
#pragma omp parallel
  {
    std::vector<double> g(kMlpWeights, 0.0);
    double part = 0.0;
#pragma omp for schedule(static)
    for (int s = 0; s < n; ++s) {
      R h[kMlpHidden + 1][kMlpFeat], z[kMlpOut];
      forward_one<R>(w, in + (size_t)s * kMlpIn, h, z);
      R y[kMlpOut], dz[kMlpOut];
      for (int o = 0; o < kMlpOut; ++o) y[o] = std::exp(z[o]);
      const R lum = (R)0.212671 * y[0] + (R)0.715160 * y[1] + (R)0.072169 * y[2];
      for (int o = 0; o < kMlpOut; ++o) {
        const R t = (R)target[(size_t)s * kMlpOut + o];
        const R d = y[o] - t;
        R scale = 1;
        if (loss == 0) scale = (R)1 / (lum * lum + (R)0.01);
        else if (loss == 1) scale = (R)1 / (y[o] * y[o] + (R)0.01);
        part += (double)(d * d * scale) * inv_n;
        // through exp, its derivative floored at exp_floor
        dz[o] = (R)2 * d * scale * (R)inv_n * std::fmax(y[o], (R)exp_floor);
      }
      // output layer
      const float* Wo = w + (size_t)kMlpHidden * kMlpFeat * kMlpFeat;
      R dh[kMlpFeat];
      for (int i = 0; i < kMlpFeat; ++i) dh[i] = 0;
      for (int o = 0; o < kMlpOut; ++o) {
        double* go = &g[(size_t)kMlpHidden * kMlpFeat * kMlpFeat + o * kMlpFeat];
        for (int i = 0; i < kMlpFeat; ++i) {
          go[i] += (double)(dz[o] * h[kMlpHidden][i]);
          dh[i] += (R)Wo[o * kMlpFeat + i] * dz[o];
        }
      }
      for (int l = kMlpHidden - 1; l >= 0; --l) {
        R dzl[kMlpFeat];
        for (int o = 0; o < kMlpFeat; ++o) dzl[o] = h[l + 1][o] > 0 ? dh[o] : (R)0;
        const float* W = w + (size_t)l * kMlpFeat * kMlpFeat;
        double* gl = &g[(size_t)l * kMlpFeat * kMlpFeat];
        R dprev[kMlpFeat];
        for (int i = 0; i < kMlpFeat; ++i) dprev[i] = 0;
        for (int o = 0; o < kMlpFeat; ++o) {
          if (dzl[o] == 0) continue;
          for (int i = 0; i < kMlpFeat; ++i) {
            gl[o * kMlpFeat + i] += (double)(dzl[o] * h[l][i]);
            dprev[i] += (R)W[o * kMlpFeat + i] * dzl[o];
          }
        }
        for (int i = 0; i < kMlpFeat; ++i) dh[i] = dprev[i];
      }
    }
#pragma omp critical
    {
      total += part;
      for (int k = 0; k < kMlpWeights; ++k) grad[k] += (R)g[k];
    }
  }
  return total;
}

template void mlp_encode<double>(const float*, double*);
template void mlp_encode<float>(const float*, float*);
template void mlp_forward<double>(const float*, const float*, int, double*);
template void mlp_forward<float>(const float*, const float*, int, float*);
template double mlp_loss_grad<double>(const float*, const float*, const float*, int, int, double, double*);
template double mlp_loss_grad<float>(const float*, const float*, const float*, int, int, double, float*);

}  // namespace orc
