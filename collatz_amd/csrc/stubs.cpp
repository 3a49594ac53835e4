// Entry points still being built; they report COLLATZ_EINVAL with a message.
#include "host_common.hpp"

using namespace cz;

extern "C" {
int collatz_server_create(const collatz_sieve*, const collatz_server_opts*, collatz_server**) {
  return fail(COLLATZ_EINVAL, "collatz_server_create: not available in this build");
}
int64_t collatz_server_submit(collatz_server*, const collatz_interval*) {
  return fail(COLLATZ_EINVAL, "collatz_server_submit: not available in this build");
}
int collatz_server_wait(collatz_server*, int64_t, double, collatz_request_result*,
                        collatz_chunk*, uint64_t) {
  return fail(COLLATZ_EINVAL, "collatz_server_wait: not available in this build");
}
int collatz_server_destroy(collatz_server*) { return COLLATZ_OK; }
}
