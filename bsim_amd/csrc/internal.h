/*
 * internal.h -- definitions shared by the host translation units of the
 * library (not part of the C ABI).
 */
#ifndef BSIM_INTERNAL_H
#define BSIM_INTERNAL_H

#include <string>

#include "network.h"

struct bsim_network {
    bsim::HostNet net;
};

/* message of the last failure on this thread (bsim_last_error) */
extern thread_local std::string g_bsim_err;

#endif
