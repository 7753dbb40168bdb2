// device.h -- device-side tables and shared helpers of the cosmo_hip library.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "internal.h"

#define HIPCHK(expr)                                                                         \
  do {                                                                                       \
    hipError_t _e = (expr);                                                                  \
    if (_e != hipSuccess)                                                                    \
      return ::cosmo::fail(COSMO_EHIP, std::string(#expr) + ": " + hipGetErrorString(_e));   \
  } while (0)

namespace cosmo {

// Hamiltonian tables as the GPU kernels read them.  All pointers are device pointers.
// Orbital g is bit g of a 64-bit state mask (neutrons [0,nn), protons [nn,nn+np)).
// Pair (a,b), a<b, has the compact index pid = b*(b-1)/2 + a, P pairs in total.
struct DevTables {
  int nn, np, norb, P;
  int nv, zv;                 // valence neutrons / protons
  int nnconf;                 // neutron configurations with a non-empty proton sector
  const uint32_t *nmask;      // [nnconf]
  const int64_t *noff;        // [nnconf+1]
  const int32_t *nsector;     // [nnconf]
  const uint32_t *pmask;      // [all proton configs, sector-major]
  const int64_t *psec_start;  // [nsec+1]
  const int32_t *pcfg_rank;   // [C(np,zv)] by colex rank
  const int32_t *pcfg_sec;    // [C(np,zv)] sector of each proton config
  const int32_t *ncfg_index;  // [C(nn,nv)] by colex rank (-1: not a basis block)
  // truncated basis (null otherwise): block b keeps the proton configurations of sector
  // ranks krank[kstart[b] .. kstart[b+1]), ascending
  const int32_t *krank;
  const int64_t *kstart;      // [nnconf+1]
  const double *vdense;       // [P*P]  V(annihilate pair pa, create pair pc)
  const uint32_t *vexist;     // [P*P/32 + 1] key present bits
  const int32_t *fstart;      // [P+1]  pure two-body terms by annihilated pair
  const int32_t *fterm;       //        created pair of each
  const int32_t *rstart;      // [P+1]  pure two-body terms by created pair
  const int32_t *rterm;       //        annihilated pair of each
  const int32_t *ob_hi_start; // [norb+1] one-body partners c > a
  const uint8_t *ob_hi;
  const int32_t *ob_lo_start; // [norb+1] one-body partners a < c
  const uint8_t *ob_lo;
  const double *esp;          // [norb*norb] one-body ESP(a -> c)
  const uint8_t *esp_exist;   // [norb*norb]
  const uint8_t *pair_a;      // [P] smaller orbital of pair pid
  const uint8_t *pair_b;      // [P] larger orbital of pair pid
};

// host-owned device allocation of the tables
struct DevTablesOwner {
  DevTables t{};
  void *blob = nullptr;  // one allocation holding every table
  size_t bytes = 0;
};
int upload_tables(const System &s, DevTablesOwner &o, hipStream_t st);
void free_tables(DevTablesOwner &o);

hipStream_t lib_stream();

// Library device memory is stream-ordered (hipMallocAsync / hipFreeAsync on the calling
// thread's library stream, from the device's pool with no release threshold): a free
// does not synchronise the device, so solves running concurrently from several host
// threads (batches, the server) do not stall one another, and memory freed by one solve
// is reused by the next.  Memory used on a foreign stream is freed only after the device
// has been synchronised (see cosmo_op_destroy).
hipError_t dmalloc_void(void **p, size_t bytes, hipStream_t st);
void dfree(void *p);
template <typename T>
inline hipError_t dmalloc(T **p, size_t bytes, hipStream_t st = nullptr) {
  return dmalloc_void(reinterpret_cast<void **>(p), bytes, st);
}
}  // namespace cosmo
