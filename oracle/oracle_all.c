/* One translation unit for the whole CPU oracle (test infrastructure only): the path-tracer and
 * denoiser restatements reuse the real-time tier's scalar functions. */
#include "rt_oracle.c"
#include "pt_oracle.c"
#include "atrous_oracle.c"

/* The ABI header's hash this library was built against (oracle/__init__.py compares it). */
const char* oracle_header_hash(void) { return MJPBR_HEADER_HASH; }
