// mlp_oracle.cc -- TEST INFRASTRUCTURE ONLY.
//
// CPU restatement, in FP64, of the pass-time cost model's network and its
// mean-squared-error gradient (qsim_amd/csrc/costmodel.hip): the checker the
// GPU forward pass and gradients are compared against.  qsim itself has no
// cost model, so there is no reference function to follow; the restatement
// is pinned by the mathematics instead (tests/test_oracle.py):
//   forward   a network with hand-set weights whose output is known in closed
//             form (identity-like layers, ReLU cut-offs);
//   gradient  central finite differences of the loss in FP64 for every
//             parameter of small networks (relative error < 1e-6), which a
//             dropped term, wrong sign or transposed weight fails.
//
// Parameter layout (as the library's): layer by layer W (in x out,
// row-major) then b; layers F -> H -> H -> H -> 1, ReLU after the hidden
// layers.

#include <cstddef>
#include <vector>

namespace {

constexpr int kLayers = 4;

struct Shape {
  unsigned in[kLayers], out[kLayers];
  size_t woff[kLayers], boff[kLayers];
  size_t n;
  Shape(unsigned F, unsigned H) {
    size_t off = 0;
    for (int l = 0; l < kLayers; ++l) {
      in[l] = l == 0 ? F : H;
      out[l] = l == kLayers - 1 ? 1 : H;
      woff[l] = off;
      off += size_t(in[l]) * out[l];
      boff[l] = off;
      off += out[l];
    }
    n = off;
  }
};

// z[l][s * out + o] = sum_i a_{l-1}[s][i] W[i][o] + b[o]; a = relu(z) for
// hidden layers, a_{-1} = x.
void Forward(const Shape& sh, const float* p, const float* x, unsigned B,
             std::vector<double> z[kLayers]) {
  for (int l = 0; l < kLayers; ++l) {
    const unsigned I = sh.in[l], O = sh.out[l];
    z[l].assign(size_t(B) * O, 0.0);
    for (unsigned s = 0; s < B; ++s) {
      for (unsigned o = 0; o < O; ++o) {
        double acc = p[sh.boff[l] + o];
        for (unsigned i = 0; i < I; ++i) {
          double a = l == 0 ? double(x[size_t(s) * I + i])
                            : z[l - 1][size_t(s) * I + i];
          if (l > 0 && a < 0) a = 0;
          acc += a * double(p[sh.woff[l] + size_t(i) * O + o]);
        }
        z[l][size_t(s) * O + o] = acc;
      }
    }
  }
}

}  // namespace

extern "C" {

unsigned long orc_mlp_num_params(unsigned F, unsigned H) {
  return Shape(F, H).n;
}

void orc_mlp_forward(unsigned F, unsigned H, const float* params,
                     const float* x, unsigned B, double* y) {
  Shape sh(F, H);
  std::vector<double> z[kLayers];
  Forward(sh, params, x, B, z);
  for (unsigned s = 0; s < B; ++s) y[s] = z[kLayers - 1][s];
}

// loss = mean_s (y_s - t_s)^2; grad = d loss / d params (FP64).
double orc_mlp_gradient(unsigned F, unsigned H, const float* params,
                        const float* x, const float* t, unsigned B,
                        double* grad) {
  Shape sh(F, H);
  std::vector<double> z[kLayers];
  Forward(sh, params, x, B, z);
  double loss = 0;
  std::vector<double> dz(B);
  for (unsigned s = 0; s < B; ++s) {
    const double d = z[kLayers - 1][s] - double(t[s]);
    loss += d * d;
    dz[s] = 2.0 * d / B;
  }
  for (size_t i = 0; i < sh.n; ++i) grad[i] = 0;
  for (int l = kLayers - 1; l >= 0; --l) {
    const unsigned I = sh.in[l], O = sh.out[l];
    std::vector<double> dprev(l > 0 ? size_t(B) * I : 0, 0.0);
    for (unsigned s = 0; s < B; ++s) {
      for (unsigned o = 0; o < O; ++o) {
        const double g = dz[size_t(s) * O + o];
        grad[sh.boff[l] + o] += g;
        for (unsigned i = 0; i < I; ++i) {
          double a = l == 0 ? double(x[size_t(s) * I + i])
                            : z[l - 1][size_t(s) * I + i];
          if (l > 0 && a < 0) a = 0;
          grad[sh.woff[l] + size_t(i) * O + o] += a * g;
          if (l > 0) {
            dprev[size_t(s) * I + i] +=
                g * double(params[sh.woff[l] + size_t(i) * O + o]);
          }
        }
      }
    }
    if (l > 0) {
      for (size_t j = 0; j < dprev.size(); ++j)
        if (z[l - 1][j] <= 0) dprev[j] = 0;
      dz.swap(dprev);
    }
  }
  return loss / B;
}

}  // extern "C"
