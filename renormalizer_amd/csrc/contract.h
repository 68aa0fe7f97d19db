// Internal interface between the contraction, Krylov and Davidson units.
#pragma once

#include "reno_common.h"

namespace reno {

struct HopLayout {
  size_t wl_bytes;  // permuted MPO (nsite 1)
  size_t t1_bytes;  // stage-1 intermediate
  size_t t2_bytes;  // stage-2 intermediate (nsite 1)
  size_t part_bytes;  // split-K partial sums
  size_t total;
};

int check_hop(const reno_hop_desc* op);
HopLayout hop_layout(const reno_hop_desc* op);
int permute_w(int prec, const void* W, void* Wl, int Dl, int d, int Dr, hipStream_t s);
// dot_part (optional): per-block partials of Re(<cout, dot_v>) over dot_np
// blocks, fused into the last stage
int hop_apply_core(const reno_hop_desc* op, const void* Wl, const void* cin, void* cout,
                   void* t1, void* t2, hipStream_t s, const int* skip, void* part,
                   size_t part_bytes, const double2* dot_v = nullptr, double* dot_part = nullptr,
                   int dot_np = 0);
size_t env_ws_bytes(int prec, int64_t Ml, int64_t d, int64_t Mr, int64_t Dl, int64_t Dr);
int env_update_core(int domain, int prec, int64_t Ml, int64_t d, int64_t Mr, int64_t Dl,
                    int64_t Dr, const void* env_in, const void* ms, const void* ms_conj,
                    const void* W, void* env_out, void* ws, hipStream_t s);

inline int64_t hop_vec_len(const reno_hop_desc* op) {
  return (int64_t)op->Ml * op->d * op->Mr;
}

}  // namespace reno
