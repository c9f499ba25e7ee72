// qsim_oracle.cc -- TEST INFRASTRUCTURE ONLY.
//
// A CPU restatement of qsim's state-vector algorithms, written from the
// reference's behaviour (files cited per function, paths relative to the
// reference checkout).  It is the checker that tests/, __graft_entry__.smoke()
// and bench.py's cpu_baseline leg compare the HIP backend against.  Nothing in
// the product (qsim_amd/, include/) links, loads or calls it.
//
// Layout: every state is interleaved complex ("normal order", the layout of
// lib/statespace_basic.h:32-37): amplitude i is (s[2i], s[2i+1]).
//
// How each function is pinned (tests/test_oracle.py, tests/test_oracle_ref.py):
//   apply_gate            known amplitudes of TestApplyGate1/2/3
//                         (tests/simulator_testfixture.h:36-200), the analytic
//                         TestMultiQubitGates closed form (:1146-1213), a dense
//                         2^n x 2^n Kronecker product at n<=8, and the compiled
//                         reference SimulatorBasic/SimulatorAVX on the grid
//                         circuits (oracle/_ref).
//   apply_controlled_gate expected_results of TestCircuitWithControlledGates
//                         (tests/simulator_testfixture.h:661-731) via the
//                         parser's "c" gates, and a dense controlled product.
//   expectation_value     <psi|U|psi> computed densely, Pauli closed forms.
//   norm / inner products closed form 0.5 + 0.5/sqrt(2)
//                         (tests/statespace_testfixture.h:466-512).
//   sample                oracle/_ref bitstrings on the same state file, and
//                         the TestSamplingSmall distribution (:514-558).
//   measure / collapse    oracle/_ref on measurement circuits, brute force.
//   generate_random_values oracle/_ref draws (same libstdc++ mt19937 and
//                         uniform_real_distribution).
//   parser / gate table   oracle/_ref final states on circuit files that use
//                         every gate name.

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <random>
#include <sstream>
#include <string>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace oracle {

// Inserts a zero bit at every set position of `mask` (mask bits processed from
// low to high), i.e. the index arithmetic of SimulatorBase::FillIndices
// (lib/simulator.h:39-66) written as a bit-deposit.
static inline uint64_t InsertZeros(uint64_t i, const unsigned* pos, unsigned k) {
  for (unsigned j = 0; j < k; ++j) {
    uint64_t low = i & ((uint64_t{1} << pos[j]) - 1);
    i = ((i >> pos[j]) << (pos[j] + 1)) | low;
  }
  return i;
}

// Parallel-deposit of the low bits of `v` into the set bits of `mask`
// (lib/bits.h:20-86, ExpandBits).
static inline uint64_t Deposit(uint64_t v, uint64_t mask) {
  uint64_t r = 0;
  for (unsigned i = 0; i < 64; ++i) {
    if ((mask >> i) & 1) {
      r |= (v & 1) << i;
      v >>= 1;
    }
  }
  return r;
}

static inline bool Parallel(unsigned n) { return n >= 14; }

// psi <- U psi on every 2^k group; lib/simulator_basic.h:167-215 (ApplyGateH).
// Products and sums in FP, matrix row-major with row bit j <-> qs[j].
template <typename FP>
void ApplyGate(unsigned n, FP* s, const unsigned* qs, unsigned k,
               const FP* m) {
  if (k == 0) {
    // One-element matrix: a global phase (lib/simulator_basic.h:167-215, H=0).
    uint64_t size = uint64_t{1} << n;
    for (uint64_t i = 0; i < size; ++i) {
      FP re = s[2 * i], im = s[2 * i + 1];
      s[2 * i] = re * m[0] - im * m[1];
      s[2 * i + 1] = re * m[1] + im * m[0];
    }
    return;
  }
  const unsigned dim = 1u << k;
  std::vector<uint64_t> offs(dim);
  for (unsigned j = 0; j < dim; ++j) {
    uint64_t a = 0;
    for (unsigned b = 0; b < k; ++b) {
      if ((j >> b) & 1) a |= uint64_t{1} << qs[b];
    }
    offs[j] = a;
  }
  int64_t ngroups = int64_t{1} << (n - k);
#pragma omp parallel for if (Parallel(n)) schedule(static)
  for (int64_t g = 0; g < ngroups; ++g) {
    FP rs[64], is[64];
    uint64_t base = InsertZeros(uint64_t(g), qs, k);
    for (unsigned j = 0; j < dim; ++j) {
      rs[j] = s[2 * (base + offs[j])];
      is[j] = s[2 * (base + offs[j]) + 1];
    }
    for (unsigned r = 0; r < dim; ++r) {
      const FP* row = m + 2 * dim * r;
      FP rn = rs[0] * row[0] - is[0] * row[1];
      FP in = rs[0] * row[1] + is[0] * row[0];
      for (unsigned c = 1; c < dim; ++c) {
        rn += rs[c] * row[2 * c] - is[c] * row[2 * c + 1];
        in += rs[c] * row[2 * c + 1] + is[c] * row[2 * c];
      }
      s[2 * (base + offs[r])] = rn;
      s[2 * (base + offs[r]) + 1] = in;
    }
  }
}

// Controlled U: the group is updated only where the control bits equal the
// deposited control values; lib/simulator_basic.h:217-275 with the masks of
// GetMasks7 (lib/simulator.h:427-441).  cvals bit i <-> cqs[i] (sorted).
template <typename FP>
void ApplyControlledGate(unsigned n, FP* s, const unsigned* qs, unsigned k,
                         const unsigned* cqs, unsigned nc, uint64_t cvals,
                         const FP* m) {
  if (nc == 0) {
    ApplyGate(n, s, qs, k, m);
    return;
  }
  uint64_t cmask = 0;
  for (unsigned i = 0; i < nc; ++i) cmask |= uint64_t{1} << cqs[i];
  uint64_t cv = Deposit(cvals, cmask);
  const unsigned dim = 1u << k;
  std::vector<uint64_t> offs(dim);
  for (unsigned j = 0; j < dim; ++j) {
    uint64_t a = 0;
    for (unsigned b = 0; b < k; ++b) {
      if ((j >> b) & 1) a |= uint64_t{1} << qs[b];
    }
    offs[j] = a;
  }
  int64_t ngroups = int64_t{1} << (n - k);
#pragma omp parallel for if (Parallel(n)) schedule(static)
  for (int64_t g = 0; g < ngroups; ++g) {
    uint64_t base = InsertZeros(uint64_t(g), qs, k);
    if ((base & cmask) != cv) continue;
    FP rs[64], is[64];
    for (unsigned j = 0; j < dim; ++j) {
      rs[j] = s[2 * (base + offs[j])];
      is[j] = s[2 * (base + offs[j]) + 1];
    }
    for (unsigned r = 0; r < dim; ++r) {
      const FP* row = m + 2 * dim * r;
      FP rn = rs[0] * row[0] - is[0] * row[1];
      FP in = rs[0] * row[1] + is[0] * row[0];
      for (unsigned c = 1; c < dim; ++c) {
        rn += rs[c] * row[2 * c] - is[c] * row[2 * c + 1];
        in += rs[c] * row[2 * c + 1] + is[c] * row[2 * c];
      }
      s[2 * (base + offs[r])] = rn;
      s[2 * (base + offs[r]) + 1] = in;
    }
  }
}

// Fixed-chunk reduction so that results do not depend on the thread count:
// chunk sums in any order, then an in-order sum of the chunk sums.
template <typename T, typename F>
T ChunkedSum(uint64_t size, F f) {
  const uint64_t chunk = uint64_t{1} << 14;
  int64_t nchunks = int64_t((size + chunk - 1) / chunk);
  std::vector<T> part(nchunks, T(0));
#pragma omp parallel for if (size >= (uint64_t{1} << 16)) schedule(static)
  for (int64_t c = 0; c < nchunks; ++c) {
    T acc = T(0);
    uint64_t i1 = std::min(size, uint64_t(c + 1) * chunk);
    for (uint64_t i = uint64_t(c) * chunk; i < i1; ++i) acc += f(i);
    part[c] = acc;
  }
  T r = T(0);
  for (auto& p : part) r += p;
  return r;
}

// <psi|U|psi>: per-group products in FP, accumulation in double;
// lib/simulator_basic.h:277-340 (ExpectationValueH).
template <typename FP>
std::complex<double> ExpectationValue(unsigned n, const FP* s,
                                      const unsigned* qs, unsigned k,
                                      const FP* m) {
  const unsigned dim = 1u << k;
  std::vector<uint64_t> offs(dim);
  for (unsigned j = 0; j < dim; ++j) {
    uint64_t a = 0;
    for (unsigned b = 0; b < k; ++b) {
      if ((j >> b) & 1) a |= uint64_t{1} << qs[b];
    }
    offs[j] = a;
  }
  auto f = [&](uint64_t g) -> std::complex<double> {
    FP rs[64], is[64];
    uint64_t base = InsertZeros(g, qs, k);
    for (unsigned j = 0; j < dim; ++j) {
      rs[j] = s[2 * (base + offs[j])];
      is[j] = s[2 * (base + offs[j]) + 1];
    }
    double re = 0, im = 0;
    for (unsigned r = 0; r < dim; ++r) {
      const FP* row = m + 2 * dim * r;
      FP rn = rs[0] * row[0] - is[0] * row[1];
      FP in = rs[0] * row[1] + is[0] * row[0];
      for (unsigned c = 1; c < dim; ++c) {
        rn += rs[c] * row[2 * c] - is[c] * row[2 * c + 1];
        in += rs[c] * row[2 * c + 1] + is[c] * row[2 * c];
      }
      re += rs[r] * rn + is[r] * in;
      im += rs[r] * in - is[r] * rn;
    }
    return {re, im};
  };
  return ChunkedSum<std::complex<double>>(uint64_t{1} << (n - k), f);
}

// lib/statespace_basic.h:155-181 (InnerProduct): sum conj(a_i) b_i in double.
template <typename FP>
std::complex<double> InnerProduct(unsigned n, const FP* a, const FP* b) {
  return ChunkedSum<std::complex<double>>(uint64_t{1} << n, [&](uint64_t i) {
    double re = a[2 * i] * b[2 * i] + a[2 * i + 1] * b[2 * i + 1];
    double im = a[2 * i] * b[2 * i + 1] - a[2 * i + 1] * b[2 * i];
    return std::complex<double>(re, im);
  });
}

// lib/statespace_basic.h:183-198 (RealInnerProduct).
template <typename FP>
double RealInnerProduct(unsigned n, const FP* a, const FP* b) {
  return ChunkedSum<double>(uint64_t{1} << n, [&](uint64_t i) -> double {
    return a[2 * i] * b[2 * i] + a[2 * i + 1] * b[2 * i + 1];
  });
}

// lib/util.h:67-85 (GenerateRandomValues): mt19937(seed), uniform [0, max),
// sorted ascending, followed by a sentinel equal to max.
static std::vector<double> GenerateRandomValues(uint64_t num, unsigned seed,
                                                double max_value) {
  std::vector<double> rs;
  rs.reserve(num + 1);
  std::mt19937 rgen(seed);
  std::uniform_real_distribution<double> distr(0.0, max_value);
  for (uint64_t i = 0; i < num; ++i) rs.push_back(distr(rgen));
  std::sort(rs.begin(), rs.end());
  rs.push_back(max_value);
  return rs;
}

// lib/statespace_basic.h:195-233 (Sample): serial FP64 norm, sorted draws,
// one serial cumulative-sum walk; draw r maps to the first k with r < csum_k;
// unmatched draws map to 2^n - 1.
template <typename FP>
void SampleWithDraws(unsigned n, const FP* p, const double* rs, uint64_t num,
                     uint64_t* out) {
  uint64_t size = uint64_t{1} << n;
  uint64_t m = 0;
  double csum = 0;
  for (uint64_t k = 0; k < size && m < num; ++k) {
    double re = p[2 * k];
    double im = p[2 * k + 1];
    csum += re * re + im * im;
    while (m < num && rs[m] < csum) out[m++] = k;
  }
  for (; m < num; ++m) out[m] = size - 1;
}

template <typename FP>
double SerialNorm(unsigned n, const FP* p) {
  double norm = 0;
  uint64_t size = uint64_t{1} << n;
  for (uint64_t k = 0; k < size; ++k) {
    double re = p[2 * k];
    double im = p[2 * k + 1];
    norm += re * re + im * im;
  }
  return norm;
}

template <typename FP>
void Sample(unsigned n, const FP* p, uint64_t num, unsigned seed,
            uint64_t* out) {
  if (num == 0) return;
  double norm = SerialNorm(n, p);
  auto rs = GenerateRandomValues(num, seed, norm);
  SampleWithDraws(n, p, rs.data(), num, out);
}

// Collapse: masked FP64 norm, then zero the non-matching amplitudes and scale
// the matching ones by 1/sqrt(norm) (lib/statespace_basic.h:237-262).
template <typename FP>
void Collapse(unsigned n, FP* s, uint64_t mask, uint64_t bits) {
  uint64_t size = uint64_t{1} << n;
  double norm = ChunkedSum<double>(size, [&](uint64_t i) -> double {
    return (i & mask) == bits
               ? double(s[2 * i] * s[2 * i] + s[2 * i + 1] * s[2 * i + 1])
               : 0.0;
  });
  FP renorm = FP(1.0 / std::sqrt(norm));
  for (uint64_t i = 0; i < size; ++i) {
    bool keep = (i & mask) == bits;
    s[2 * i] = keep ? s[2 * i] * renorm : 0;
    s[2 * i + 1] = keep ? s[2 * i + 1] * renorm : 0;
  }
}

// One-partition VirtualMeasure + Collapse (lib/statespace.h:84-140 with the
// FindMeasuredBits walk of lib/statespace_basic.h:275-295).  The draw is
// RandomValue(rgen, norm) (lib/util.h:61-65).
template <typename FP>
bool Measure(unsigned n, FP* s, const unsigned* qubits, unsigned nq,
             std::mt19937& rgen, uint64_t* bits_out) {
  uint64_t mask = 0;
  for (unsigned i = 0; i < nq; ++i) {
    if (qubits[i] >= n) return false;
    mask |= uint64_t{1} << qubits[i];
  }
  uint64_t size = uint64_t{1} << n;
  double norm = ChunkedSum<double>(size, [&](uint64_t i) -> double {
    return s[2 * i] * s[2 * i] + s[2 * i + 1] * s[2 * i + 1];
  });
  std::uniform_real_distribution<double> distr(0.0, norm);
  double r = distr(rgen);
  double csum = 0;
  uint64_t found = size - 1;
  for (uint64_t k = 0; k < size; ++k) {
    FP re = s[2 * k], im = s[2 * k + 1];
    csum += re * re + im * im;
    if (r < csum) {
      found = k;
      break;
    }
  }
  uint64_t bits = found & mask;
  Collapse(n, s, mask, bits);
  *bits_out = bits;
  return true;
}

// ---------------------------------------------------------------------------
// Gate table of the qsim text format (lib/gates_qsim.h, lib/gate.h:266-290,
// lib/circuit_qsim_parser.h:264-385).  Matrices are evaluated in FP exactly as
// the reference's templates do (constants cast to FP, angles as FP).

template <typename FP>
struct Op {
  int kind;  // 0 gate, 1 measurement
  unsigned time;
  std::vector<unsigned> qubits;      // sorted
  std::vector<unsigned> controls;    // sorted
  uint64_t cmask;                    // control values, bit i <-> controls[i]
  std::vector<FP> matrix;
};

// Permute matrix indices so that it acts on ascending qubits;
// lib/matrix.h:230-293 (NormalToGateOrderPermutation + MatrixShuffle).
template <typename FP>
static void SortQubits(std::vector<unsigned>& qs, std::vector<FP>& m,
                       bool symmetric) {
  unsigned q = qs.size();
  bool sorted = true;
  for (unsigned i = 1; i < q; ++i)
    if (qs[i] < qs[i - 1]) sorted = false;
  if (sorted) return;
  std::vector<unsigned> idx(q);
  for (unsigned i = 0; i < q; ++i) idx[i] = i;
  std::sort(idx.begin(), idx.end(),
            [&](unsigned a, unsigned b) { return qs[a] < qs[b]; });
  if (!symmetric) {
    unsigned dim = 1u << q;
    std::vector<FP> mt = m;
    auto perm_bits = [&](unsigned x) {
      unsigned r = 0;
      for (unsigned i = 0; i < q; ++i) r |= ((x >> i) & 1) << idx[i];
      return r;
    };
    for (unsigned i = 0; i < dim; ++i) {
      unsigned pi = perm_bits(i);
      for (unsigned j = 0; j < dim; ++j) {
        unsigned pj = perm_bits(j);
        m[2 * (dim * i + j)] = mt[2 * (dim * pi + pj)];
        m[2 * (dim * i + j) + 1] = mt[2 * (dim * pi + pj) + 1];
      }
    }
  }
  std::sort(qs.begin(), qs.end());
}

template <typename FP>
static bool MakeGate(const std::string& name, const std::vector<FP>& prm,
                     std::vector<unsigned> qs, Op<FP>& op) {
  const FP h = static_cast<FP>(0.5);
  const FP is2 = static_cast<FP>(0.7071067811865475);
  std::vector<FP> m;
  bool symmetric = true;
  auto need = [&](size_t nq, size_t np) {
    return qs.size() == nq && prm.size() == np;
  };
  if (name == "p") {  // GateGPh, gates_qsim.h:68-84
    if (!need(0, 1)) return false;
    FP phi = prm[0];
    m = {FP(std::cos(phi)), FP(std::sin(phi))};
  } else if (name == "id1") {  // :90-100
    if (!need(1, 0)) return false;
    m = {1, 0, 0, 0, 0, 0, 1, 0};
  } else if (name == "h") {  // :106-118
    if (!need(1, 0)) return false;
    m = {is2, 0, is2, 0, is2, 0, -is2, 0};
  } else if (name == "t") {  // :124-136
    if (!need(1, 0)) return false;
    m = {1, 0, 0, 0, 0, 0, is2, is2};
  } else if (name == "x") {  // :142-152
    if (!need(1, 0)) return false;
    m = {0, 0, 1, 0, 1, 0, 0, 0};
  } else if (name == "y") {  // :158-168
    if (!need(1, 0)) return false;
    m = {0, 0, 0, -1, 0, 1, 0, 0};
  } else if (name == "z") {  // :174-184
    if (!need(1, 0)) return false;
    m = {1, 0, 0, 0, 0, 0, -1, 0};
  } else if (name == "x_1_2") {  // :190-202
    if (!need(1, 0)) return false;
    m = {h, h, h, -h, h, -h, h, h};
  } else if (name == "y_1_2") {  // :208-220
    if (!need(1, 0)) return false;
    m = {h, h, -h, -h, h, h, h, h};
  } else if (name == "rx" || name == "ry" || name == "rz") {  // :227-283
    if (!need(1, 1)) return false;
    FP phi2 = -0.5 * prm[0];
    FP c = std::cos(phi2);
    FP s = std::sin(phi2);
    if (name == "rx") m = {c, 0, 0, s, 0, s, c, 0};
    if (name == "ry") m = {c, 0, s, 0, -s, 0, c, 0};
    if (name == "rz") m = {c, s, 0, 0, 0, 0, c, -s};
  } else if (name == "rxy") {  // :289-306
    if (!need(1, 2)) return false;
    FP theta = prm[0];
    FP phi2 = -0.5 * prm[1];
    FP cp = std::cos(phi2);
    FP sp = std::sin(phi2);
    FP ct = std::cos(theta) * sp;
    FP st = std::sin(theta) * sp;
    m = {cp, 0, st, ct, -st, ct, cp, 0};
  } else if (name == "hz_1_2") {  // :312-326
    if (!need(1, 0)) return false;
    m = {h, h, 0, -is2, is2, 0, h, h};
  } else if (name == "s") {  // :332-342
    if (!need(1, 0)) return false;
    m = {1, 0, 0, 0, 0, 0, 0, 1};
  } else if (name == "id2") {  // :368-387
    if (!need(2, 0)) return false;
    m = {1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0,
         0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 0};
  } else if (name == "cz") {  // :393-413
    if (!need(2, 0)) return false;
    m = {1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0,
         0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, -1, 0};
  } else if (name == "cnot" || name == "cx") {  // :419-440, not symmetric
    if (!need(2, 0)) return false;
    symmetric = false;
    m = {1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 0,
         0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0};
  } else if (name == "sw") {  // :446-470
    if (!need(2, 0)) return false;
    m = {1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 0, 0, 0,
         0, 0, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 0};
  } else if (name == "is") {  // :476-503
    if (!need(2, 0)) return false;
    m = {1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 0, 0,
         0, 0, 0, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 0};
  } else if (name == "fs") {  // :509-533
    if (!need(2, 2)) return false;
    FP theta = prm[0];
    FP phi = prm[1];
    if (phi < 0) phi += 2 * 3.141592653589793;
    FP ct = std::cos(theta);
    FP st = std::sin(theta);
    FP cp = std::cos(phi);
    FP sp = std::sin(phi);
    m = {1, 0, 0, 0, 0, 0, 0, 0, 0, 0, ct, 0, 0, -st, 0, 0,
         0, 0, 0, -st, ct, 0, 0, 0, 0, 0, 0, 0, 0, 0, cp, -sp};
  } else if (name == "cp") {  // :582-600
    if (!need(2, 1)) return false;
    FP cp = std::cos(prm[0]);
    FP sp = std::sin(prm[0]);
    m = {1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0,
         0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, cp, -sp};
  } else {
    return false;
  }
  if (qs.size() == 2 && qs[0] == qs[1]) return false;
  SortQubits(qs, m, symmetric);
  op.kind = 0;
  op.qubits = qs;
  op.matrix = m;
  op.cmask = 0;
  return true;
}

static unsigned GateArity(const std::string& name, unsigned* nparams) {
  static const struct {
    const char* name;
    unsigned nq, np;
  } table[] = {{"p", 0, 1},     {"id1", 1, 0},   {"h", 1, 0},
               {"t", 1, 0},     {"x", 1, 0},     {"y", 1, 0},
               {"z", 1, 0},     {"x_1_2", 1, 0}, {"y_1_2", 1, 0},
               {"rx", 1, 1},    {"ry", 1, 1},    {"rz", 1, 1},
               {"rxy", 1, 2},   {"hz_1_2", 1, 0}, {"s", 1, 0},
               {"id2", 2, 0},   {"cz", 2, 0},    {"cnot", 2, 0},
               {"cx", 2, 0},    {"sw", 2, 0},    {"is", 2, 0},
               {"fs", 2, 2},    {"cp", 2, 1}};
  for (auto& e : table) {
    if (name == e.name) {
      *nparams = e.np;
      return e.nq;
    }
  }
  *nparams = 0;
  return 99;
}

template <typename FP>
static FP ParseNumber(const std::string& tok, bool* ok);

template <>
float ParseNumber<float>(const std::string& tok, bool* ok) {
  char* end;
  float v = std::strtof(tok.c_str(), &end);
  *ok = end != tok.c_str() && *end == 0;
  return v;
}

template <>
double ParseNumber<double>(const std::string& tok, bool* ok) {
  char* end;
  double v = std::strtod(tok.c_str(), &end);
  *ok = end != tok.c_str() && *end == 0;
  return v;
}

template <typename FP>
struct Circuit {
  unsigned num_qubits = 0;
  std::vector<Op<FP>> ops;
};

// The qsim text format (docs/input_format.md; lib/circuit_qsim_parser.h:52-447):
// first non-comment line = qubit count, then "time name qubits params" per line,
// "c q.. name .." for controlled gates, "m q.." for measurements.
template <typename FP>
static bool ParseCircuit(const std::string& text, unsigned maxtime,
                         Circuit<FP>& c) {
  std::stringstream fs(text);
  std::string line;
  c.num_qubits = 0;
  c.ops.clear();
  while (std::getline(fs, line)) {
    if (line.empty() || line[0] == '#') continue;
    std::stringstream ss(line);
    std::vector<std::string> toks;
    std::string t;
    while (ss >> t) toks.push_back(t);
    if (toks.empty()) continue;
    if (c.num_qubits == 0) {
      c.num_qubits = std::atoi(toks[0].c_str());
      if (c.num_qubits == 0) return false;
      continue;
    }
    if (toks.size() < 2) return false;
    unsigned time = std::atoi(toks[0].c_str());
    if (time > maxtime) break;
    size_t pos = 1;
    std::vector<unsigned> controls;
    if (toks[1] == "c") {
      pos = 2;
      while (pos < toks.size() && std::isdigit((unsigned char)toks[pos][0])) {
        controls.push_back(std::atoi(toks[pos].c_str()));
        ++pos;
      }
      if (controls.empty() || pos >= toks.size()) return false;
    }
    std::string name = toks[pos++];
    Op<FP> op;
    op.time = time;
    if (name == "m") {
      if (!controls.empty()) return false;
      op.kind = 1;
      for (; pos < toks.size(); ++pos) {
        unsigned q = std::atoi(toks[pos].c_str());
        if (q >= c.num_qubits) return false;
        op.qubits.push_back(q);
      }
      op.cmask = 0;
      c.ops.push_back(op);
      continue;
    }
    unsigned np;
    unsigned nq = GateArity(name, &np);
    if (nq == 99 || toks.size() != pos + nq + np) return false;
    std::vector<unsigned> qs;
    for (unsigned i = 0; i < nq; ++i) {
      unsigned q = std::atoi(toks[pos + i].c_str());
      if (q >= c.num_qubits) return false;
      qs.push_back(q);
    }
    std::vector<FP> prm;
    for (unsigned i = 0; i < np; ++i) {
      bool ok;
      prm.push_back(ParseNumber<FP>(toks[pos + nq + i], &ok));
      if (!ok) return false;
    }
    if (!MakeGate<FP>(name, prm, qs, op)) return false;
    if (!controls.empty()) {
      // MakeControlledGate with all control values 1
      // (lib/gate.h:36-70, lib/circuit_qsim_parser.h:387-447).
      std::sort(controls.begin(), controls.end());
      for (auto cq : controls) {
        if (cq >= c.num_qubits) return false;
        for (auto q : op.qubits)
          if (q == cq) return false;
      }
      op.controls = controls;
      op.cmask = (uint64_t{1} << controls.size()) - 1;
    }
    c.ops.push_back(op);
  }
  return c.num_qubits > 0;
}

template <typename FP>
static bool RunCircuit(const Circuit<FP>& c, FP* s, unsigned seed,
                       std::vector<uint64_t>* mresults) {
  std::mt19937 rgen(seed);
  for (const auto& op : c.ops) {
    if (op.kind == 1) {
      uint64_t bits;
      if (!Measure(c.num_qubits, s, op.qubits.data(), op.qubits.size(), rgen,
                   &bits))
        return false;
      if (mresults) mresults->push_back(bits);
    } else {
      ApplyControlledGate(c.num_qubits, s, op.qubits.data(), op.qubits.size(),
                          op.controls.data(), op.controls.size(), op.cmask,
                          op.matrix.data());
    }
  }
  return true;
}

}  // namespace oracle

// ---------------------------------------------------------------------------
// C ABI for ctypes (tests and the bench's CPU leg).

using namespace oracle;

extern "C" {

int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void orc_set_num_threads(int n) {
#ifdef _OPENMP
  omp_set_num_threads(n);
#else
  (void)n;
#endif
}

#define ORC_FP_FUNCS(SUF, FP)                                                  \
  void orc_apply_gate_##SUF(unsigned n, FP* s, const unsigned* qs, unsigned k, \
                            const FP* m) {                                     \
    ApplyGate<FP>(n, s, qs, k, m);                                             \
  }                                                                            \
  void orc_apply_controlled_gate_##SUF(unsigned n, FP* s, const unsigned* qs,  \
                                       unsigned k, const unsigned* cqs,        \
                                       unsigned nc, uint64_t cvals,            \
                                       const FP* m) {                          \
    ApplyControlledGate<FP>(n, s, qs, k, cqs, nc, cvals, m);                   \
  }                                                                            \
  void orc_expectation_value_##SUF(unsigned n, const FP* s,                    \
                                   const unsigned* qs, unsigned k,             \
                                   const FP* m, double* out) {                 \
    auto r = ExpectationValue<FP>(n, s, qs, k, m);                             \
    out[0] = r.real();                                                         \
    out[1] = r.imag();                                                         \
  }                                                                            \
  double orc_norm_##SUF(unsigned n, const FP* s) {                             \
    return RealInnerProduct<FP>(n, s, s);                                      \
  }                                                                            \
  double orc_serial_norm_##SUF(unsigned n, const FP* s) {                      \
    return SerialNorm<FP>(n, s);                                               \
  }                                                                            \
  void orc_inner_product_##SUF(unsigned n, const FP* a, const FP* b,           \
                               double* out) {                                  \
    auto r = InnerProduct<FP>(n, a, b);                                        \
    out[0] = r.real();                                                         \
    out[1] = r.imag();                                                         \
  }                                                                            \
  double orc_real_inner_product_##SUF(unsigned n, const FP* a, const FP* b) {  \
    return RealInnerProduct<FP>(n, a, b);                                      \
  }                                                                            \
  void orc_sample_##SUF(unsigned n, const FP* s, uint64_t num, unsigned seed,  \
                        uint64_t* out) {                                       \
    Sample<FP>(n, s, num, seed, out);                                          \
  }                                                                            \
  void orc_sample_with_draws_##SUF(unsigned n, const FP* s, const double* rs,  \
                                   uint64_t num, uint64_t* out) {              \
    SampleWithDraws<FP>(n, s, rs, num, out);                                   \
  }                                                                            \
  void orc_collapse_##SUF(unsigned n, FP* s, uint64_t mask, uint64_t bits) {   \
    Collapse<FP>(n, s, mask, bits);                                            \
  }                                                                            \
  void* orc_circuit_parse_##SUF(const char* text, unsigned maxtime) {          \
    auto* c = new Circuit<FP>();                                               \
    if (!ParseCircuit<FP>(text, maxtime, *c)) {                                \
      delete c;                                                                \
      return nullptr;                                                          \
    }                                                                          \
    return c;                                                                  \
  }                                                                            \
  unsigned orc_circuit_num_qubits_##SUF(void* h) {                             \
    return static_cast<Circuit<FP>*>(h)->num_qubits;                           \
  }                                                                            \
  unsigned orc_circuit_num_ops_##SUF(void* h) {                                \
    return static_cast<Circuit<FP>*>(h)->ops.size();                           \
  }                                                                            \
  void orc_circuit_free_##SUF(void* h) { delete static_cast<Circuit<FP>*>(h); }\
  int orc_circuit_run_##SUF(void* h, FP* s, unsigned seed,                     \
                            uint64_t* mres, unsigned max_mres,                 \
                            unsigned* num_mres) {                              \
    std::vector<uint64_t> r;                                                   \
    bool ok = RunCircuit<FP>(*static_cast<Circuit<FP>*>(h), s, seed, &r);      \
    unsigned cnt = std::min<unsigned>(r.size(), max_mres);                     \
    for (unsigned i = 0; i < cnt; ++i) mres[i] = r[i];                         \
    if (num_mres) *num_mres = r.size();                                        \
    return ok ? 0 : -1;                                                        \
  }                                                                            \
  int orc_gate_matrix_##SUF(const char* name, const FP* prm, unsigned np,      \
                            const unsigned* qs, unsigned nq,                   \
                            unsigned* qs_out, FP* m_out) {                     \
    Op<FP> op;                                                                 \
    std::vector<FP> p(prm, prm + np);                                          \
    std::vector<unsigned> q(qs, qs + nq);                                      \
    if (!MakeGate<FP>(name, p, q, op)) return -1;                              \
    for (unsigned i = 0; i < op.qubits.size(); ++i) qs_out[i] = op.qubits[i];  \
    for (unsigned i = 0; i < op.matrix.size(); ++i) m_out[i] = op.matrix[i];   \
    return 0;                                                                  \
  }

ORC_FP_FUNCS(f32, float)
ORC_FP_FUNCS(f64, double)

void orc_generate_random_values(uint64_t num, unsigned seed, double max_value,
                                double* out) {
  auto rs = GenerateRandomValues(num, seed, max_value);
  std::memcpy(out, rs.data(), (num + 1) * sizeof(double));
}

}  // extern "C"
