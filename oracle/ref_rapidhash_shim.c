/*
 * Build shim for oracle/_ref: exposes the reference's own vendored rapidhash
 * mixer (include/sqsgen/core/helpers/rapidhash.h:149-212, compiled from its
 * location in the reference tree) so the oracle's restatement can be checked
 * against it.  Test infrastructure only.
 */
#include <stdint.h>
#include "rapidhash.h"

uint64_t ref_rapid_mix(uint64_t a, uint64_t b) { return rapid_mix(a, b); }
uint64_t ref_rapid_secret(int i) { return rapid_secret[i & 7]; }
