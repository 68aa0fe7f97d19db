// Device Householder QR (householder.hip), used by the blocked svd_qn QR/RQ.
#pragma once

#include "reno_common.h"

namespace reno {

// workspace bytes for qr_householder(m, n)
size_t qr_ws_bytes(int m, int n);
// true when the register-resident panel covers m rows
bool qr_supported(int m, int n);
// In place QR of X (m x n, column-major, ld = m): on return the upper
// triangle of X holds R (k x n, k = min(m, n), real diagonal) and Q (m x k,
// column-major, ld = m) holds the orthonormal factor.
int qr_householder(hipStream_t s, int m, int n, double2* X, double2* Q, void* ws);

}  // namespace reno
