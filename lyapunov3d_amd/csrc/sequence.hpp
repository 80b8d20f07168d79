/*
 * sequence.hpp - the letter sequence's host-side check (no HIP in it).
 */
#pragma once

#include <stdint.h>

#include "lyap.h"

namespace lyap {

/* Period of a -1-terminated sequence of letter indices 0..3 (A..D) held in
 * an array of LYAP_MAX_SEQ + 1 entries; 0 when the sequence is empty, longer
 * than LYAP_MAX_SEQ or holds any other value. */
inline int sequence_period(const int32_t *seq)
{
    int L = 0;
    while (L <= LYAP_MAX_SEQ && seq[L] != -1) {
        if (seq[L] < 0 || seq[L] > 3) return 0;
        ++L;
    }
    return L > LYAP_MAX_SEQ ? 0 : L;
}

} /* namespace lyap */
