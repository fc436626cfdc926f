// Device-wide integer primitives on the library stream, written for wave64:
// exclusive prefix sums of ints and stable LSD radix sorts of (key, value)
// pairs. Temporary storage is the caller's (sized by the *TempBytes
// functions, 256-byte aligned), so callers keep their own allocation pattern.
#pragma once

#include <stddef.h>

namespace cuhc {

// out[i] = in[0] + ... + in[i - 1] for i < count (out[0] = 0). in may equal out.
size_t scanTempBytes(int count);
int exclusiveScan(const int *in, int *out, int count, void *temp);

// Stable sort of count (key, value) pairs by the low `bits` bits of the keys
// (higher key bits must be zero), keysIn/valuesIn -> keysOut/valuesOut
// (inputs left unchanged). One 8-bit digit per pass, least significant first.
size_t sortTempBytes32(int count);
int radixSortPairs(const unsigned int *keysIn, unsigned int *keysOut, const int *valuesIn,
                   int *valuesOut, int count, int bits, void *temp);
size_t sortTempBytes64(int count);
int radixSortPairs(const unsigned long long *keysIn, unsigned long long *keysOut,
                   const unsigned int *valuesIn, unsigned int *valuesOut, int count, int bits,
                   void *temp);

} // namespace cuhc
