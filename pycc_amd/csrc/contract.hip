// Einsum-style contraction planner and strided elementwise kernels (see contract.h).
#include "contract.h"
#include "split16.h"
#include <type_traits>
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>

namespace ccx {

TV tv(const void* p, const char* lab, std::vector<int64_t> ext) {
  TV v;
  v.p = p;
  v.lab = lab;
  CCX_CHECK((int)ext.size() == v.rank(), ERR_INTERNAL, std::string("view rank mismatch: ") + lab);
  v.ext = ext;
  v.str.assign(ext.size(), 1);
  for (int i = (int)ext.size() - 2; i >= 0; --i) v.str[i] = v.str[i + 1] * ext[i + 1];
  return v;
}

TV tvs(const void* p, const char* lab, std::vector<int64_t> ext, std::vector<int64_t> str) {
  TV v;
  v.p = p;
  v.lab = lab;
  CCX_CHECK((int)ext.size() == v.rank() && str.size() == ext.size(), ERR_INTERNAL,
            std::string("view rank mismatch: ") + lab);
  v.ext = ext;
  v.str = str;
  return v;
}

namespace {

// CCX_TRACE=1 prints every planned contraction to stderr (diagnostics only).
bool trace_on() {
  static int on = -1;
  if (on < 0) {
    const char* e = std::getenv("CCX_TRACE");
    on = (e && e[0] == '1') ? 1 : 0;
  }
  return on == 1;
}

// One index of a GEMM class: its extent and its stride in each of the (up to 3) operands.
struct Lab {
  char c;
  int64_t ext;
  int64_t s[3];   // stride in A, B, C (0 when absent)
};

// Fuse labels whose strides nest in all operands in `ops` (bitmask over A=1, B=2, C=4).
void fuse(std::vector<Lab>& g, int ops) {
  // drop unit extents
  g.erase(std::remove_if(g.begin(), g.end(), [](const Lab& l) { return l.ext == 1; }), g.end());
  bool changed = true;
  while (changed && g.size() > 1) {
    changed = false;
    for (size_t x = 0; x < g.size() && !changed; ++x)
      for (size_t y = 0; y < g.size() && !changed; ++y) {
        if (x == y) continue;
        bool ok = true;
        for (int t = 0; t < 3; ++t)
          if ((ops >> t) & 1)
            if (g[x].s[t] != g[y].ext * g[y].s[t]) ok = false;
        if (ok) {
          Lab m = g[y];
          m.ext = g[x].ext * g[y].ext;
          g[y] = m;
          g.erase(g.begin() + x);
          changed = true;
        }
      }
  }
}

// Build the 2-level index of class `g` for operand t; `inner` selects which label is level 2.
Idx2 make_idx(const std::vector<Lab>& g, int inner, int t) {
  Idx2 d;
  if (g.empty()) { d.n2 = 1; d.s1 = 0; d.s2 = 0; return d; }
  if (g.size() == 1) { d.n2 = g[0].ext; d.s1 = 0; d.s2 = g[0].s[t]; return d; }
  int outer = 1 - inner;
  d.n2 = g[inner].ext;
  d.s1 = g[outer].s[t];
  d.s2 = g[inner].s[t];
  return d;
}

int64_t extent(const std::vector<Lab>& g) {
  int64_t e = 1;
  for (auto& l : g) e *= l.ext;
  return e;
}

// Prefer as level-2 index the label that is contiguous in operand `pref1`, then `pref2`.
int pick_inner(const std::vector<Lab>& g, int pref1, int pref2) {
  if (g.size() < 2) return 0;
  for (int t : {pref1, pref2})
    for (int i = 0; i < 2; ++i)
      if (g[i].s[t] == 1) return i;
  return std::abs(g[0].s[pref1]) <= std::abs(g[1].s[pref1]) ? 0 : 1;
}

template <typename T>
TV slice(const TV& v, char c, int64_t i) {
  TV r = v;
  int d = v.find(c);
  if (d < 0) return r;
  r.p = (const void*)((const T*)v.p + i * v.str[d]);
  r.lab.erase(d, 1);
  r.ext.erase(r.ext.begin() + d);
  r.str.erase(r.str.begin() + d);
  return r;
}

}  // namespace

template <typename T>
void contract(T alpha, const TV& A_, const TV& B_, T beta, const TV& C, Workspace& ws, hipStream_t s, bool split16) {
  // Keep C's contiguous index in the N class (B operand): the epilogue stores 16 consecutive n
  // per MFMA row, so this makes the stores coalesced.
  bool swap = false;
  for (int d = 0; d < C.rank(); ++d)
    if (C.str[d] == 1 && C.ext[d] > 1) swap = A_.find(C.lab[d]) >= 0 && B_.find(C.lab[d]) < 0;
  const TV& A = swap ? B_ : A_;
  const TV& B = swap ? A_ : B_;
  std::vector<Lab> gm, gn, gk, gb;
  auto ext_of = [&](char c) -> int64_t {
    int64_t e = -1;
    for (const TV* v : {&A, &B, &C}) {
      int d = v->find(c);
      if (d >= 0) {
        CCX_CHECK(e < 0 || e == v->ext[d], ERR_ARG, std::string("extent mismatch for index ") + c);
        e = v->ext[d];
      }
    }
    return e;
  };
  for (const TV* v : {&A, &B, &C})
    for (int i = 0; i < v->rank(); ++i)
      for (int j = i + 1; j < v->rank(); ++j)
        CCX_CHECK(v->lab[i] != v->lab[j], ERR_ARG, "repeated index in one operand is not supported");
  for (char c : C.lab) {
    int a = A.find(c), b = B.find(c);
    Lab l{c, ext_of(c), {a >= 0 ? A.str[a] : 0, b >= 0 ? B.str[b] : 0, C.str[C.find(c)]}};
    if (a >= 0 && b >= 0) gb.push_back(l);
    else if (a >= 0) gm.push_back(l);
    else if (b >= 0) gn.push_back(l);
    else CCX_CHECK(false, ERR_ARG, std::string("output index absent from inputs: ") + c);
  }
  for (char c : A.lab) {
    if (C.find(c) >= 0) continue;
    int b = B.find(c);
    CCX_CHECK(b >= 0, ERR_ARG, std::string("index summed in one operand only: ") + c);
    gk.push_back(Lab{c, ext_of(c), {A.str[A.find(c)], B.str[b], 0}});
  }
  for (char c : B.lab)
    CCX_CHECK(A.find(c) >= 0 || C.find(c) >= 0, ERR_ARG, std::string("index summed in one operand only: ") + c);

  const std::vector<Lab> om = gm, on = gn, ok = gk, ob = gb;   // before fusion
  fuse(gm, 1 | 4);
  fuse(gn, 2 | 4);
  fuse(gk, 1 | 2);
  fuse(gb, 1 | 2 | 4);

  // Too many indices left in one class: loop on the host over that class's original index with
  // the smallest extent (slicing removes it; the remainder is planned again).
  auto smallest = [](const std::vector<Lab>& g) {
    const Lab* best = nullptr;
    for (auto& l : g)
      if (l.ext > 1 && (!best || l.ext < best->ext)) best = &l;
    return best->c;
  };
  auto loop_over = [&](char c) {
    int64_t e = ext_of(c);
    for (int64_t i = 0; i < e; ++i)
      contract<T>(alpha, slice<T>(A, c, i), slice<T>(B, c, i), beta, slice<T>(C, c, i), ws, s, split16);
  };
  // An M (N) class with a third index and no batch index: that index becomes the batch, with
  // stride 0 in B (A), which lacks it. Pick the index whose removal leaves a fusable class.
  auto try_batch = [&](std::vector<Lab>& g, const std::vector<Lab>& orig, int ops, int absent) -> bool {
    if (!gb.empty() || g.size() <= 2) return false;
    for (size_t x = 0; x < orig.size(); ++x) {
      if (orig[x].ext == 1) continue;
      std::vector<Lab> rest;
      for (size_t y = 0; y < orig.size(); ++y)
        if (y != x) rest.push_back(orig[y]);
      fuse(rest, ops);
      if (rest.size() <= 2) {
        Lab bl = orig[x];
        bl.s[absent] = 0;
        gb.push_back(bl);
        g = rest;
        return true;
      }
    }
    return false;
  };
  try_batch(gm, om, 1 | 4, 1);
  try_batch(gn, on, 2 | 4, 0);
  if (trace_on()) {
    auto str = [](const std::vector<Lab>& g) {
      std::string r;
      for (auto& l : g) r += l.c;
      return r.empty() ? std::string("-") : r;
    };
    std::fprintf(stderr, "[ccx] %s,%s->%s  M{%s} N{%s} K{%s} B{%s} M=%lld N=%lld K=%lld%s\n", A_.lab.c_str(), B_.lab.c_str(),
                 C.lab.c_str(), str(gm).c_str(), str(gn).c_str(), str(gk).c_str(), str(gb).c_str(), (long long)extent(gm), (long long)extent(gn), (long long)extent(gk),
                 (gb.size() > 1 || gk.size() > 2 || gm.size() > 2 || gn.size() > 2) ? "  HOST-LOOP" : "");
  }
  if (gb.size() > 1) { loop_over(smallest(ob)); return; }
  if (gk.size() > 2) {
    // a K index is sliced by accumulating: the first slice applies beta, the rest add
    char c = smallest(ok);
    int64_t e = ext_of(c);
    for (int64_t i = 0; i < e; ++i)
      contract<T>(alpha, slice<T>(A, c, i), slice<T>(B, c, i), i == 0 ? beta : T(1), C, ws, s, split16);
    return;
  }
  if (gm.size() > 2) { loop_over(smallest(om)); return; }
  if (gn.size() > 2) { loop_over(smallest(on)); return; }

  GemmParams<T> p;
  p.M = extent(gm);
  p.N = extent(gn);
  p.K = extent(gk);
  p.batch = gb.empty() ? 1 : gb[0].ext;
  const int im = pick_inner(gm, 0, 2), in = pick_inner(gn, 1, 2), ik = pick_inner(gk, 0, 1);
  p.am = make_idx(gm, im, 0);
  p.cm = make_idx(gm, im, 2);
  p.bn = make_idx(gn, in, 1);
  p.cn = make_idx(gn, in, 2);
  p.ak = make_idx(gk, ik, 0);
  p.bk = make_idx(gk, ik, 1);
  if (!gb.empty()) { p.a_b = gb[0].s[0]; p.b_b = gb[0].s[1]; p.c_b = gb[0].s[2]; }
  p.alpha = alpha;
  p.beta = beta;
  p.A = (const T*)A.p;
  p.B = (const T*)B.p;
  p.C = (T*)C.p;
  if constexpr (std::is_same<T, double>::value) {
    if (split16 && p.batch == 1 && 2.0 * p.M * p.N * p.K >= X3_ROUTE_FLOPS && x3_contract(p, ws, s)) return;
  }
  tc_gemm<T>(p, ws, s);
}

// ---- strided elementwise ----
struct Perm6 {
  int rank;
  int64_t ext[6];
  int64_t sa[6], sc[6];
  int64_t total;
};

template <typename T>
__global__ void axpby_kernel(Perm6 d, T alpha, const T* __restrict__ A, T beta, T* __restrict__ C) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < d.total;
       e += (int64_t)gridDim.x * blockDim.x) {
    int64_t r = e, oa = 0, oc = 0;
    for (int k = d.rank - 1; k >= 0; --k) {
      int64_t q = r / d.ext[k];
      int64_t i = r - q * d.ext[k];
      r = q;
      oa += i * d.sa[k];
      oc += i * d.sc[k];
    }
    T v = A ? alpha * A[oa] : T(0);
    if (beta != T(0)) v += beta * C[oc];
    C[oc] = v;
  }
}

template <typename T>
static void launch_perm(const Perm6& d, T alpha, const T* A, T beta, T* C, hipStream_t s) {
  if (d.total == 0) return;
  int blocks = (int)std::min<int64_t>(ceil_div(d.total, 256), 16384);
  axpby_kernel<T><<<blocks, 256, 0, s>>>(d, alpha, A, beta, C);
  CCX_HIP(hipGetLastError());
}

template <typename T>
void axpby(T alpha, const TV& A, T beta, const TV& C, hipStream_t s) {
  CCX_CHECK(A.rank() == C.rank() && C.rank() <= 6, ERR_ARG, "axpby rank");
  Perm6 d{};
  d.rank = C.rank();
  d.total = 1;
  for (int k = 0; k < C.rank(); ++k) {
    int a = A.find(C.lab[k]);
    CCX_CHECK(a >= 0 && A.ext[a] == C.ext[k], ERR_ARG, "axpby labels/extents");
    d.ext[k] = C.ext[k];
    d.sc[k] = C.str[k];
    d.sa[k] = A.str[a];
    d.total *= C.ext[k];
  }
  launch_perm<T>(d, alpha, (const T*)A.p, beta, (T*)C.p, s);
}

template <typename T>
void scale(T beta, const TV& C, hipStream_t s) {
  Perm6 d{};
  d.rank = C.rank();
  d.total = 1;
  for (int k = 0; k < C.rank(); ++k) {
    d.ext[k] = C.ext[k];
    d.sc[k] = C.str[k];
    d.sa[k] = 0;
    d.total *= C.ext[k];
  }
  launch_perm<T>(d, T(0), (const T*)nullptr, beta, (T*)C.p, s);
}

template void contract<double>(double, const TV&, const TV&, double, const TV&, Workspace&, hipStream_t, bool);
template void contract<float>(float, const TV&, const TV&, float, const TV&, Workspace&, hipStream_t, bool);
template void axpby<double>(double, const TV&, double, const TV&, hipStream_t);
template void axpby<float>(float, const TV&, float, const TV&, hipStream_t);
template void scale<double>(double, const TV&, hipStream_t);
template void scale<float>(float, const TV&, hipStream_t);

}  // namespace ccx
