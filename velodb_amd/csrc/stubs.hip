// stubs.hip -- intentionally empty: every operator of include/vo_api.h is
// implemented (sort*.hip, compact.hip, join.hip, shuffle.hip, misc.hip,
// queue.hip, serve.hip).
#include "vo_internal.hpp"
