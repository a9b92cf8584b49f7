// status.h -- absl::Status-shaped error plumbing for the C ABI.
//
// The reference returns absl::Status / absl::StatusOr everywhere on the path
// (clustering/status_macros.h:26-60).  The library keeps the same codes and
// carries the message in a thread-local string readable through
// pcc_last_error().
#ifndef PCC_AMD_STATUS_H_
#define PCC_AMD_STATUS_H_

#include <string>

#include "pcc_amd.h"

namespace pcc {

int SetError(int code, const std::string& msg);
void ClearError();
const char* LastError();

// RETURN_IF_ERROR for integer status codes.
#define PCC_RETURN_IF_ERROR(expr)    \
  do {                               \
    const int _pcc_rc = (expr);      \
    if (_pcc_rc != PCC_OK) return _pcc_rc; \
  } while (0)

}  // namespace pcc

#endif  // PCC_AMD_STATUS_H_
