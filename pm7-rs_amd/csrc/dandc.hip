// SPDX-License-Identifier: GPL-3.0-or-later
//
// Divide-and-conquer SCF (Dixon-Merz; pm7-rs src/dandc/mod.rs): the system-wide kernels of one
// step. The per-subsystem work (Fock build, eigensolve, weighted local density) is in scf.hip;
// these kernels move data between the subsystems and the global density, which is stored as
// atom-pair blocks (CSR over atoms, pm7-rs src/dandc/sparse.rs):
//
//   k_dc_pop          atom populations (diagonal-block traces)
//   k_dc_global_field EV sum_{b != a} q_b / r_ab for every atom (nuclear once, electronic per step)
//   k_dc_local_field  each subsystem atom's environment: the global potential less its own cluster
//   k_dc_gather       the global blocks of every subsystem's atom pairs into its dense local matrix
//   k_dc_fermi        the common chemical potential, bracketed over every weighted state
//   k_dc_assemble     one wave per stored block: weighted sum of the subsystem densities that hold
//                     it, normalised, the RMS change and the linear mixing fused
//   k_dc_finish       RMS change and electronic energy
//
// No atomics on the global density: every element has one owner and every sum a fixed order, so a
// step gives the same bits on every run.
#include <hip/hip_runtime.h>

#include "kernels.h"
#include "pm7_device.h"
#include "wg.h"

namespace pm7 {

// Population of each atom: the trace of its diagonal block (SparseDensity::population,
// src/dandc/sparse.rs).
__global__ void __launch_bounds__(256) k_dc_pop(int n, const int64_t* __restrict__ diag_off,
                                                const int* __restrict__ norb, const double* __restrict__ P,
                                                double* __restrict__ pop) {
  const int a = blockIdx.x * 256 + threadIdx.x;
  if (a >= n) return;
  const int no = norb[a];
  const double* d = P + diag_off[a];
  double s = 0.0;
  for (int k = 0; k < no; ++k) s += d[k * no + k];
  pop[a] = s;
}

// out[a] = EV sum_{b != a} q_b / |r_b - r_a|, b ascending (the molecular far field and its
// electronic half, src/dandc/mod.rs:874-958). Positions and charges stream through LDS tiles.
__global__ void __launch_bounds__(256) k_dc_global_field(int n, const double* __restrict__ xyz,
                                                         const double* __restrict__ q, double* __restrict__ out) {
  __shared__ double s_x[256], s_y[256], s_z[256], s_q[256];
  const int a = blockIdx.x * 256 + threadIdx.x;
  double xa = 0.0, ya = 0.0, za = 0.0;
  if (a < n) {
    xa = xyz[3 * a];
    ya = xyz[3 * a + 1];
    za = xyz[3 * a + 2];
  }
  double s = 0.0;
  for (int b0 = 0; b0 < n; b0 += 256) {
    const int b = b0 + threadIdx.x;
    __syncthreads();
    if (b < n) {
      s_x[threadIdx.x] = xyz[3 * b];
      s_y[threadIdx.x] = xyz[3 * b + 1];
      s_z[threadIdx.x] = xyz[3 * b + 2];
      s_q[threadIdx.x] = q[b];
    }
    __syncthreads();
    const int nb = min(256, n - b0);
    if (a < n)
      for (int t = 0; t < nb; ++t) {
        if (b0 + t == a) continue;
        const double dx = s_x[t] - xa, dy = s_y[t] - ya, dz = s_z[t] - za;
        s += EV * s_q[t] / sqrt(dx * dx + dy * dy + dz * dz);
      }
  }
  if (a < n) out[a] = s;
}

// Environment of batch atom t (subsystem atom i of subsystem mi): g[parent] less the monopoles of
// the subsystem's own other atoms, images closer than 1e-8 Bohr skipped (environment_field,
// src/dandc/mod.rs:961-1011).
__global__ void __launch_bounds__(256) k_dc_local_field(int n_batch_atoms, const int* __restrict__ atom_mol,
                                                        const MolDev* __restrict__ mols,
                                                        const int* __restrict__ parent,
                                                        const double* __restrict__ xyz,
                                                        const double* __restrict__ q, const double* __restrict__ g,
                                                        double* __restrict__ out) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= n_batch_atoms) return;
  const MolDev& M = mols[atom_mol[t]];
  const double xa = xyz[3 * t], ya = xyz[3 * t + 1], za = xyz[3 * t + 2];
  double s = 0.0;
  for (int j = M.atom0; j < M.atom0 + M.n_atoms; ++j) {
    if (j == t) continue;
    const double dx = xyz[3 * j] - xa, dy = xyz[3 * j + 1] - ya, dz = xyz[3 * j + 2] - za;
    const double r = sqrt(dx * dx + dy * dy + dz * dz);
    if (r < 1.0e-8) continue;
    s += EV * q[parent[j]] / r;
  }
  out[t] = g[parent[t]] - s;
}

// H' = H - diag(v_nuc): the nuclear half of the environment, constant over the SCF
// (src/dandc/mod.rs:545-552). One workgroup per subsystem.
__global__ void __launch_bounds__(256) k_dc_shift_h(const MolDev* __restrict__ mols, const int* __restrict__ atom_z,
                                                    const int* __restrict__ atom_ao,
                                                    const ElemDev* __restrict__ elem, double* __restrict__ slab,
                                                    const double* __restrict__ v_nuc) {
  const MolDev M = mols[blockIdx.x];
  double* H = slab + M.mat0 + (int64_t)S_H * M.ld * M.ld;
  for (int la = threadIdx.x; la < M.n_atoms; la += 256) {
    const int a = M.atom0 + la;
    const int o = atom_ao[a], no = elem[atom_z[a]].n_orb;
    for (int k = 0; k < no; ++k) H[(int64_t)(o + k) * M.ld + o + k] -= v_nuc[a];
  }
}

// Every subsystem's dense local density from the global blocks of its atom pairs (local_density,
// src/dandc/mod.rs:628-645); a pair the global pattern does not hold reads as zero. UHF: the total
// into DU_PT and P_alpha = (P + S) / 2, P_beta = (P - S) / 2 into DU_PA, DU_PB
// (src/dandc/mod.rs:568-588).
__global__ void __launch_bounds__(256) k_dc_gather(const MolDev* __restrict__ mols, const int* __restrict__ aocode,
                                                   const int64_t* __restrict__ ao0,
                                                   const int64_t* __restrict__ amap,
                                                   const int64_t* __restrict__ amap0,
                                                   const double* __restrict__ P, const double* __restrict__ S,
                                                   double* __restrict__ slab) {
  const int mi = blockIdx.x;
  const MolDev M = mols[mi];
  const int n = M.nao, ld = M.ld, na = M.n_atoms;
  const int* code = aocode + ao0[mi];
  const int64_t* map = amap + amap0[mi];
  const int64_t sq = (int64_t)ld * ld;
  double* base = slab + M.mat0;
  for (int idx = threadIdx.x; idx < n * n; idx += 256) {
    const int mu = idx / n, nu = idx - mu * n;
    const int ca = code[mu], cb = code[nu];
    const int la = ca & 0xffff, ma = (ca >> 16) & 0xff;
    const int lb = cb & 0xffff, mb = (cb >> 16) & 0xff, nb = cb >> 24;
    const int64_t gofs = map[(int64_t)la * na + lb];
    const int64_t x = (int64_t)mu * ld + nu;
    const double p = gofs < 0 ? 0.0 : P[gofs + ma * nb + mb];
    if (!M.uhf) {
      base[DR_P * sq + x] = p;
    } else {
      const double s = gofs < 0 ? 0.0 : S[gofs + ma * nb + mb];
      base[DU_PT * sq + x] = p;
      base[DU_PA * sq + x] = 0.5 * (p + s);
      base[DU_PB * sq + x] = 0.5 * (p - s);
    }
  }
}

// Electrons at chemical potential mu over every weighted state (occupation dropped below 1e-14).
__device__ inline double dc_occ_weighted(double e, double w, double mu, double width) {
  const double x = (e - mu) / width;
  const double f = x > 40.0 ? 0.0 : (x < -40.0 ? 2.0 : 2.0 / (1.0 + exp(x)));
  return f >= 1.0e-14 ? f * w : 0.0;
}

constexpr int FERMI_SPLIT = 15;   // interior points per narrowing step (the bracket shrinks 16x)
constexpr int FERMI_BLOCKS = 128;  // workgroups of one count pass (the states split between them)
constexpr int FERMI_BT = 256;
constexpr int FERMI_STEPS = 14;    // 16^13 > 1e15: the 1000 eV bracket is below 1e-12 eV after 13 passes
static_assert(DC_FERMI_PART_WORDS == FERMI_BLOCKS * 16, "Fermi partial-count buffer");

// The common chemical potential (find_fermi, src/dandc/mod.rs:715-770). The electron count is
// monotone in mu, and the reference brackets its root in [-500, 500] eV until the bracket is below
// 1e-12 eV. Here each pass evaluates the count at 15 interior points of the bracket at once and
// the next pass keeps the sub-interval where it crosses n_elec: 13 passes instead of 50 halvings,
// the same root to within the 1e-12 eV bracket. A pass is one launch of FERMI_BLOCKS workgroups
// over the states (partial sums per workgroup, fixed order); the next launch's workgroups each
// reduce those partials in workgroup order and take the same decision. Pass 0 counts at -500 and
// 500 eV: if that bracket does not hold n_elec electrons, scal[3] = 1 and the host refuses the
// step. The last launch writes scal[2] = E_F. fst[2][4]: bracket, failed, done (ping-pong);
// fpart[2][FERMI_BLOCKS][16]: the partial counts (ping-pong).
__global__ void __launch_bounds__(FERMI_BT) k_dc_fermi_step(int step, int64_t n_states, const double* __restrict__ eps,
                                                            const double* __restrict__ wn, double n_elec,
                                                            double width, double* __restrict__ fpart,
                                                            double* __restrict__ fst, double* __restrict__ scal) {
  __shared__ double red[(Wg<FERMI_BT>::NW + 1) * FERMI_SPLIT];
  double lo = -500.0, hi = 500.0;
  bool failed = false, done = false;
  if (step > 0) {
    const double* ps = fst + 4 * ((step - 1) & 1);
    lo = ps[0];
    hi = ps[1];
    failed = ps[2] != 0.0;
    done = ps[3] != 0.0;
    if (!failed && !done) {
      const double* pp = fpart + (int64_t)((step - 1) & 1) * FERMI_BLOCKS * 16;
      double c[FERMI_SPLIT];
#pragma unroll
      for (int j = 0; j < FERMI_SPLIT; ++j) {
        double t = 0.0;
        if (step > 1 || j < 2)
          for (int b = 0; b < FERMI_BLOCKS; ++b) t += pp[b * 16 + j];
        c[j] = t;
      }
      if (step == 1) {
        failed = c[0] > n_elec || c[1] < n_elec;
      } else {
        // The first interior point holding n_elec electrons closes the new bracket.
        const double stp = (hi - lo) / (FERMI_SPLIT + 1);
        double nlo = lo, nhi = hi;
        bool found = false;
#pragma unroll
        for (int j = 0; j < FERMI_SPLIT; ++j) {
          const double mu = lo + (j + 1) * stp;
          if (!found) {
            if (c[j] < n_elec) {
              nlo = mu;
            } else {
              nhi = mu;
              found = true;
            }
          }
        }
        lo = nlo;
        hi = nhi;
        done = hi - lo < 1.0e-12;
      }
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    double* ws = fst + 4 * (step & 1);
    ws[0] = lo;
    ws[1] = hi;
    ws[2] = failed ? 1.0 : 0.0;
    ws[3] = done ? 1.0 : 0.0;
    if (step == FERMI_STEPS || failed || done) {
      scal[2] = failed ? 0.0 : 0.5 * (lo + hi);
      scal[3] = failed ? 1.0 : 0.0;
    }
  }
  if (failed || done || step == FERMI_STEPS) return;  // the same decision in every workgroup
  // This pass's partial counts: at lo and hi (pass 0) or at the 15 interior points.
  const double stp = (hi - lo) / (FERMI_SPLIT + 1);
  double v[FERMI_SPLIT];
#pragma unroll
  for (int j = 0; j < FERMI_SPLIT; ++j) v[j] = 0.0;
  for (int64_t k = (int64_t)blockIdx.x * FERMI_BT + threadIdx.x; k < n_states; k += (int64_t)FERMI_BLOCKS * FERMI_BT) {
    const double w = wn[k];
    if (w == 0.0) continue;
    const double e = eps[k];
    if (step == 0) {
      v[0] += dc_occ_weighted(e, w, lo, width);
      v[1] += dc_occ_weighted(e, w, hi, width);
    } else if ((e - lo) / width < -40.0) {
      // saturated over the whole bracket (rounding is monotone, so at every interior point too)
#pragma unroll
      for (int j = 0; j < FERMI_SPLIT; ++j) v[j] += 2.0 * w;
    } else if (!((e - hi) / width > 40.0)) {
#pragma unroll
      for (int j = 0; j < FERMI_SPLIT; ++j) v[j] += dc_occ_weighted(e, w, lo + (j + 1) * stp, width);
    }
  }
  wg_sum_n<FERMI_BT, FERMI_SPLIT>(v, red);
  double* pw = fpart + (int64_t)(step & 1) * FERMI_BLOCKS * 16 + blockIdx.x * 16;
  if (threadIdx.x == 0)
#pragma unroll
    for (int j = 0; j < FERMI_SPLIT; ++j) pw[j] = v[j];
}

// One wave per stored atom-pair block: the next density is the partition-weighted sum of the
// subsystem densities that hold the block, times the block's normaliser (the inverse of its total
// weight, src/dandc/mod.rs:232-258, 330-398); the RMS change against the current density is
// accumulated (rms_diff, src/dandc/sparse.rs) and the current density is mixed in place,
// P <- (1 - m) P + m P_next (src/dandc/sparse.rs mix; src/dandc/mod.rs:400-407). UHF: the spin
// density from the same blocks, alpha minus beta channel. partial[slot] = the block's sum of
// squared changes.
__global__ void __launch_bounds__(256) k_dc_assemble(int64_t n_slots, const int64_t* __restrict__ slot_off,
                                                     const int* __restrict__ slot_dims,
                                                     const int64_t* __restrict__ cstart,
                                                     const DcContrib* __restrict__ contrib,
                                                     const double* __restrict__ normalizer,
                                                     const double* __restrict__ slab, double* __restrict__ P,
                                                     double* __restrict__ S, double mixing, int uhf,
                                                     double* __restrict__ partial) {
  const int64_t slot = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (slot >= n_slots) return;
  const int dims = slot_dims[slot];
  const int na = dims & 0xff, nb = dims >> 8;
  const int64_t c0 = cstart[slot], c1 = cstart[slot + 1];
  const double nrm = normalizer[slot];
  double* p = P + slot_off[slot];
  double* s = uhf ? S + slot_off[slot] : nullptr;
  double sq = 0.0;
  for (int e = lane; e < na * nb; e += 64) {
    const int mu = e / nb, nu = e - mu * nb;
    double acc_a = 0.0, acc_b = 0.0;
    for (int64_t c = c0; c < c1; ++c) {
      const DcContrib& q = contrib[c];
      const int64_t x = q.base + (int64_t)mu * q.ld + nu;
      acc_a += q.w * slab[x];
      if (uhf) acc_b += q.w * slab[x + q.dstep];
    }
    const double next = (uhf ? acc_a + acc_b : acc_a) * nrm;
    const double old = p[e];
    const double d = old - next;
    sq += d * d;
    p[e] = (1.0 - mixing) * old + mixing * next;
    if (uhf) s[e] = (1.0 - mixing) * s[e] + mixing * ((acc_a - acc_b) * nrm);
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) sq += __shfl_down(sq, off, 64);
  if (lane == 0) partial[slot] = sq;
}

// scal[0] = RMS change over the stored elements, scal[1] = electronic energy (sum of the
// subsystems' weighted contributions, src/dandc/mod.rs:408).
__global__ void __launch_bounds__(1024) k_dc_finish(int64_t n_slots, const double* __restrict__ partial,
                                                    int n_sub, const double* __restrict__ e_part, int64_t n_el,
                                                    double* __restrict__ scal) {
  __shared__ double red[(Wg<1024>::NW + 1) * 2];
  double v[2] = {0.0, 0.0};
  for (int64_t k = threadIdx.x; k < n_slots; k += 1024) v[0] += partial[k];
  for (int k = threadIdx.x; k < n_sub; k += 1024) v[1] += e_part[k];
  wg_sum_n<1024, 2>(v, red);
  if (threadIdx.x == 0) {
    scal[0] = sqrt(v[0] / (double)(n_el > 0 ? n_el : 1));
    scal[1] = v[1];
  }
}

hipError_t launch_dc_pop(const DcDev& d, hipStream_t s) {
  hipLaunchKernelGGL(k_dc_pop, dim3((d.n_atoms + 255) / 256), dim3(256), 0, s, d.n_atoms, d.diag_off, d.norb,
                     d.P, d.pop);
  return hipGetLastError();
}

hipError_t launch_dc_field(const BatchDev& b, const DcDev& d, const double* q, double* g, double* v,
                           hipStream_t s) {
  hipLaunchKernelGGL(k_dc_global_field, dim3((d.n_atoms + 255) / 256), dim3(256), 0, s, d.n_atoms, d.xyz, q, g);
  hipLaunchKernelGGL(k_dc_local_field, dim3((d.n_batch_atoms + 255) / 256), dim3(256), 0, s, d.n_batch_atoms,
                     d.atom_mol, b.mols, d.parent, b.xyz, q, g, v);
  return hipGetLastError();
}

hipError_t launch_dc_shift_h(const BatchDev& b, const DcDev& d, hipStream_t s) {
  hipLaunchKernelGGL(k_dc_shift_h, dim3(d.n_sub), dim3(256), 0, s, b.mols, b.atom_z, b.atom_ao, b.elem, b.slab,
                     d.v_nuc);
  return hipGetLastError();
}

hipError_t launch_dc_gather(const BatchDev& b, const DcDev& d, hipStream_t s) {
  hipLaunchKernelGGL(k_dc_gather, dim3(d.n_sub), dim3(256), 0, s, b.mols, d.aocode, d.ao0, d.amap, d.amap0, d.P,
                     d.S, b.slab);
  return hipGetLastError();
}

hipError_t launch_dc_fermi(const BatchDev& b, const DcDev& d, double n_elec, double width, hipStream_t s) {
  for (int step = 0; step <= FERMI_STEPS; ++step)
    hipLaunchKernelGGL(k_dc_fermi_step, dim3(step == FERMI_STEPS ? 1 : FERMI_BLOCKS), dim3(FERMI_BT), 0, s, step,
                       d.n_states, b.vecs, d.wnorm, n_elec, width, d.fermi_part, d.fermi_state, d.scal);
  return hipGetLastError();
}

hipError_t launch_dc_assemble(const BatchDev& b, const DcDev& d, double mixing, hipStream_t s) {
  if (d.n_slots <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_dc_assemble, dim3((unsigned)((d.n_slots + 3) / 4)), dim3(256), 0, s, d.n_slots, d.slot_off,
                     d.slot_dims, d.cstart, d.contrib, d.normalizer, b.slab, d.P, d.S, mixing, d.uhf, d.partial);
  return hipGetLastError();
}

hipError_t launch_dc_finish(const DcDev& d, hipStream_t s) {
  hipLaunchKernelGGL(k_dc_finish, dim3(1), dim3(1024), 0, s, d.n_slots, d.partial, d.n_sub, d.e_part, d.n_el,
                     d.scal);
  return hipGetLastError();
}

}  // namespace pm7
