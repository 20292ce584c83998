// Curvature-driven knot refinement, host side as in the reference:
// knotsToAdd (nurbs.h:255-281), findSpan A2.1 (nurbs.h:35-49), multiplicity
// (nurbs.h:283-287), curveKnotInsert A5.1 (nurbs.h:97-140) and
// curvatureRefinementU/V (nurbs.h:290-486).  FP32 in the reference's
// operation order (compiled with -ffp-contract=off).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/rtnurbs.h"

namespace {

struct P3 { float x, y, z; };
struct P4 { float x, y, z, w; };

int find_span(int n, int p, float u, const float* U) {
  if (u == U[n + 1]) return n;
  int low = p, high = n + 1, mid = (low + high) / 2;
  while (u < U[mid] || u >= U[mid + 1]) {
    if (u < U[mid]) high = mid; else low = mid;
    mid = (low + high) / 2;
  }
  return mid;
}

int knot_multiplicity(const float* U, int k, float u) {
  int s = 0;
  for (int i = k; i >= 0 && U[i] == u; --i) ++s;
  return s;
}

float length3(P3 a) { return std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }

int knots_to_add(const P3* P, const float* U, int i, float C, int p) {
  float vsum = 0;
  P3 Vj[RTN_MAX_DEGREE + 1];
  const int base = i - p + 1;
  for (int j = i - p + 1; j <= i; j++) {
    const float d = U[j + p] - U[j];
    if (d > 0.0f) {
      const float s = (float)p / d;
      Vj[j - base] = {(P[j].x - P[j - 1].x) * s, (P[j].y - P[j - 1].y) * s, (P[j].z - P[j - 1].z) * s};
    } else {
      Vj[j - base] = {0, 0, 0};
    }
    vsum += length3(Vj[j - base]);
  }
  float amod = 0;
  for (int j = i - p + 2; j <= i; j++) {
    const float d = U[j + p - 1] - U[j];
    P3 A{0, 0, 0};
    if (d > 0.0f) {
      const float s = (float)(p - 1) / d;
      A = {(Vj[j - base].x - Vj[j - 1 - base].x) * s, (Vj[j - base].y - Vj[j - 1 - base].y) * s,
           (Vj[j - base].z - Vj[j - 1 - base].z) * s};
    }
    amod = fmaxf(amod, length3(A));
  }
  const float num = amod * powf(U[i + 1] - U[i], 1.5f);
  const float den = sqrtf(vsum / p);
  if (!(den > 1e-9f)) return 0;
  const float nf = C * num / den;
  if (!(nf > 1e-9f)) return 0;
  return (int)nf;
}

// A5.1 for one insertion (r = 1).  Returns false where the reference prints
// an error and leaves its outputs unset (r + s > p).
bool insert_once(int np, int p, const float* UP, float u, int k, int s, float* UQ, const P4* Pw, P4* Qw) {
  const int r = 1;
  if (r + s > p) return false;
  const int mp = np + p + 1;
  int L = 0;
  P4 Rw[RTN_MAX_DEGREE + 1];
  for (int i = 0; i <= k; i++) UQ[i] = UP[i];
  for (int i = 1; i <= r; i++) UQ[k + i] = u;
  for (int i = k + 1; i <= mp; i++) UQ[i + r] = UP[i];
  for (int i = 0; i <= k - p; i++) Qw[i] = Pw[i];
  for (int i = k - s; i <= np; i++) Qw[i + r] = Pw[i];
  for (int i = 0; i <= p - s; i++) Rw[i] = Pw[k - p + i];
  for (int j = 1; j <= r; j++) {
    L = k - p + j;
    for (int i = 0; i <= p - j - s; i++) {
      const float alpha = (u - UP[L + i]) / (UP[i + k + 1] - UP[L + i]);
      const float om = 1.0f - alpha;
      Rw[i] = {Rw[i + 1].x * alpha + Rw[i].x * om, Rw[i + 1].y * alpha + Rw[i].y * om,
               Rw[i + 1].z * alpha + Rw[i].z * om, Rw[i + 1].w * alpha + Rw[i].w * om};
    }
    Qw[L] = Rw[0];
    Qw[k + r - j - s] = Rw[p - j - s];
  }
  for (int i = L + 1; i < k - s; i++) Qw[i] = Rw[i - L];
  return true;
}

}  // namespace

extern "C" int rtn_refine(int dir, int n, int p, const float* U, int m, const float* Pw, float C, int* inserted,
                          float* knots_out, float* Pw_out) {
  if (!U || !Pw || !inserted || (dir != 0 && dir != 1) || p < 1 || p > RTN_MAX_DEGREE) return RTN_EINVAL;
  const P4* P = reinterpret_cast<const P4*>(Pw);
  const int nn = dir == 0 ? n : m;     // last index along the refined direction
  const int other = dir == 0 ? m : n;  // last index along the other direction
  auto at = [&](int along, int c) -> const P4& {
    return dir == 0 ? P[(size_t)along * (m + 1) + c] : P[(size_t)c * (m + 1) + along];
  };
  std::vector<int> add(nn - p + 1);
  std::vector<P3> curve(nn + 1);
  int total = 0;
  for (int i = p; i <= nn; i++) {
    int ni = 0;
    for (int c = 0; c <= other; c++) {
      for (int a = 0; a <= nn; a++) {
        const P4& w = at(a, c);
        curve[a] = {w.x / w.w, w.y / w.w, w.z / w.w};
      }
      ni = std::max(ni, knots_to_add(curve.data(), U, i, C, p));
    }
    add[i - p] = ni;
    total += ni;
  }
  *inserted = total;
  if (!knots_out) return RTN_OK;
  if (!Pw_out) return RTN_EINVAL;
  const int nfin = nn + total;
  std::vector<float> UA(nfin + p + 2), UB(nfin + p + 2);
  std::memcpy(UA.data(), U, (nn + p + 2) * sizeof(float));
  std::vector<P4> A((size_t)(other + 1) * (nfin + 1)), B((size_t)(other + 1) * (nfin + 1));
  for (int c = 0; c <= other; c++)
    for (int a = 0; a <= nn; a++) A[(size_t)c * (nfin + 1) + a] = at(a, c);
  float *Us = UA.data(), *Ud = UB.data();
  P4 *src = A.data(), *dst = B.data();
  std::vector<P4> cin(nfin + 1), cout_(nfin + 1);
  int ncur = nn;
  for (int i = p; i <= nn; i++) {
    const int cnt = add[i - p];
    if (cnt <= 0) continue;
    const float lo = U[i], hi = U[i + 1];
    for (int j = 1; j <= cnt; j++) {
      const float u = lo + (hi - lo) * j / (float)(cnt + 1);
      const int k = find_span(ncur, p, u, Us);
      const int s = knot_multiplicity(Us, k, u);
      for (int c = 0; c <= other; c++) {
        for (int a = 0; a <= ncur; a++) cin[a] = src[(size_t)c * (nfin + 1) + a];
        if (!insert_once(ncur, p, Us, u, k, s, Ud, cin.data(), cout_.data())) return RTN_EINVAL;
        for (int a = 0; a <= ncur + 1; a++) dst[(size_t)c * (nfin + 1) + a] = cout_[a];
      }
      std::swap(src, dst);
      std::swap(Us, Ud);
      ncur++;
    }
  }
  std::memcpy(knots_out, Us, (nfin + p + 2) * sizeof(float));
  P4* O = reinterpret_cast<P4*>(Pw_out);
  for (int c = 0; c <= other; c++)
    for (int a = 0; a <= nfin; a++) {
      if (dir == 0) O[(size_t)a * (m + 1) + c] = src[(size_t)c * (nfin + 1) + a];
      else O[(size_t)c * (nfin + 1) + a] = src[(size_t)c * (nfin + 1) + a];
    }
  return RTN_OK;
}
