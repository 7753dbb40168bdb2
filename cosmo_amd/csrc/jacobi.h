// jacobi.h -- parallel cyclic Jacobi eigensolver for a small symmetric matrix held by one
// workgroup in LDS (round-robin pairing, one barrier pair per round), warm-started from the
// previous eigenvectors.  Used by the single-launch Davidson (dav_wg.h).
#pragma once

// Eigenpairs of S[0..k) (row stride LIM): eigenvalues ascending in sh.eigs (all k),
// sh.order[i] = the column of V holding the i-th; V is warm-started from the previous
// k_old = sh.kv vectors.  SH provides S (array or pointer), LDS arrays A, V (LIM x LIM),
// W2 (2 LIM x LIM), jc, js (LIM / 2), partner, pairidx, order, eigs (LIM) and the scalars
// kv, rot, dctl.
template <int LIM, class SH>
__device__ void lds_jacobi(SH &sh, int k) {
  const int tid = threadIdx.x;
  const int kv = sh.kv;
  double *A = sh.A, *V = sh.V, *A2 = sh.W2, *V2 = sh.W2 + LIM * LIM;
  // Vext = [[V_old, 0], [0, I]]  (into V2), T = S Vext (into A2), A = Vext^T T
  for (int o = tid; o < k * k; o += blockDim.x) {
    const int r = o / k, c = o % k;
    V2[r * LIM + c] = (r < kv && c < kv) ? V[r * LIM + c] : (r == c ? 1.0 : 0.0);
  }
  __syncthreads();
  for (int o = tid; o < k * k; o += blockDim.x) {
    const int r = o / k, c = o % k;
    double t = 0.0;
    for (int l = 0; l < k; ++l) t += sh.S[r * LIM + l] * V2[l * LIM + c];
    A2[r * LIM + c] = t;
  }
  __syncthreads();
  for (int o = tid; o < k * k; o += blockDim.x) {
    const int r = o / k, c = o % k;
    double t = 0.0;
    for (int l = 0; l < k; ++l) t += V2[l * LIM + r] * A2[l * LIM + c];
    A[r * LIM + c] = t;
  }
  __syncthreads();
  for (int o = tid; o < k * k; o += blockDim.x) V[(o / k) * LIM + o % k] = V2[(o / k) * LIM + o % k];
  __syncthreads();
  // symmetrise (the two products round differently in the two triangles)
  for (int o = tid; o < k * k; o += blockDim.x) {
    const int r = o / k, c = o % k;
    if (r < c) {
      const double m = 0.5 * (A[r * LIM + c] + A[c * LIM + r]);
      A[r * LIM + c] = m;
      A[c * LIM + r] = m;
    }
  }
  __syncthreads();
  // cyclic Jacobi, round-robin pairing: player K-1 meets player r, players (r+i, r-i) mod
  // K-1 meet for i = 1..K/2-1; a player >= k (odd k) sits the round out.  An element is
  // rotated away while it exceeds 1e-15 of the matrix scale (the largest |A_ii|): below
  // that it is rounding noise, which a threshold relative to small diagonal entries would
  // chase sweep after sweep.
  if (tid == 0) {
    double sc = 0.0;
    for (int i = 0; i < k; ++i) sc = fmax(sc, fabs(A[i * LIM + i]));
    for (int i = 0; i < k; ++i)
      for (int j = i + 1; j < k; ++j) sc = fmax(sc, fabs(A[i * LIM + j]));
    sh.dctl = 1e-15 * sc + 1e-300;
  }
  __syncthreads();
  const double jtol = sh.dctl;
  const int K = k + (k & 1), H = K / 2;
  for (int sweep = 0; sweep < 40 && k > 1; ++sweep) {
    // stop before a sweep that would rotate nothing: one pass over the off-diagonal part
    if (tid == 0) sh.rot = 0;
    __syncthreads();
    for (int o = tid; o < k * k; o += blockDim.x) {
      const int r = o / k, c = o % k;
      if (r < c && fabs(A[r * LIM + c]) > jtol) sh.rot = 1;
    }
    __syncthreads();
    if (!sh.rot) break;
    __syncthreads();
    for (int rr = 0; rr < K - 1; ++rr) {
      // the round's pairs and their rotations
      for (int i = tid; i < H; i += blockDim.x) {
        int p, q;
        if (i == 0) {
          p = K - 1;
          q = rr;
        } else {
          p = (rr + i) % (K - 1);
          q = (rr - i + (K - 1)) % (K - 1);
        }
        if (p > q) {
          const int t = p;
          p = q;
          q = t;
        }
        double c = 1.0, sn = 0.0;
        if (q < k) {
          const double apq = A[p * LIM + q], app = A[p * LIM + p], aqq = A[q * LIM + q];
          if (fabs(apq) > jtol) {
            // rotation zeroing A[p][q]: tan = t solves t^2 + 2 t theta - 1 = 0, smaller root
            const double theta = (aqq - app) / (2.0 * apq);
            const double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
            c = 1.0 / sqrt(t * t + 1.0);
            sn = t * c;
          }
        }
        sh.partner[i] = p;  // pair i: players p < q (q >= k: p sits the round out)
        sh.pairidx[i] = q;
        sh.jc[i] = c;
        sh.js[i] = sn;
      }
      __syncthreads();
      // A = J^T A J and V = V J with J[p][p] = J[q][q] = c, J[p][q] = s, J[q][p] = -s, in
      // place: the pairs split A into disjoint 2 x 2 blocks (rows of pair i, columns of pair
      // j) and V into disjoint row pieces; a block whose pairs both keep c = 1, s = 0 stays
      for (int o = tid; o < H * H + k * H; o += blockDim.x) {
        if (o < H * H) {
          const int i = o / H, j = o % H;
          const double si = sh.js[i], sj = sh.js[j];
          if (si == 0.0 && sj == 0.0) continue;
          const double ci = sh.jc[i], cj = sh.jc[j];
          const int pi = sh.partner[i], qi = sh.pairidx[i], pj = sh.partner[j], qj = sh.pairidx[j];
          const bool hi = qi < k, hj = qj < k;
          const double a_pp = A[pi * LIM + pj], a_pq = hj ? A[pi * LIM + qj] : 0.0;
          const double a_qp = hi ? A[qi * LIM + pj] : 0.0, a_qq = (hi && hj) ? A[qi * LIM + qj] : 0.0;
          // right by J_j, then left by J_i^T
          const double b_pp = cj * a_pp - sj * a_pq, b_pq = sj * a_pp + cj * a_pq;
          const double b_qp = cj * a_qp - sj * a_qq, b_qq = sj * a_qp + cj * a_qq;
          A[pi * LIM + pj] = ci * b_pp - si * b_qp;
          if (hj) A[pi * LIM + qj] = ci * b_pq - si * b_qq;
          if (hi) A[qi * LIM + pj] = si * b_pp + ci * b_qp;
          if (hi && hj) A[qi * LIM + qj] = si * b_pq + ci * b_qq;
        } else {
          const int x = (o - H * H) / H, j = (o - H * H) % H;
          const double sj = sh.js[j];
          if (sj == 0.0) continue;
          const double cj = sh.jc[j];
          const int pj = sh.partner[j], qj = sh.pairidx[j];  // sj != 0: both players present
          const double vp = V[x * LIM + pj], vq = V[x * LIM + qj];
          V[x * LIM + pj] = cj * vp - sj * vq;
          V[x * LIM + qj] = sj * vp + cj * vq;
        }
      }
      __syncthreads();
    }
  }
  // ascending order, lower column first on ties: column e goes to its rank
  for (int e = tid; e < k; e += blockDim.x) {
    const double ve = A[e * LIM + e];
    int rank = 0;
    for (int j = 0; j < k; ++j) {
      const double vj = A[j * LIM + j];
      rank += (vj < ve || (vj == ve && j < e)) ? 1 : 0;
    }
    sh.order[rank] = e;
    sh.eigs[rank] = ve;
  }
  if (tid == 0) sh.kv = k;
  __syncthreads();
}

