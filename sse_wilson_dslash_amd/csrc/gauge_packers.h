// The reference's C gauge packers (lib/packer_nopad.c), shared by the GPU and
// CPU builds of the drop-in API.  Include after sse_dslash.h, inside
// extern "C".
#pragma once

#include <string.h>

// lib/packer_nopad.c:29-45: u is [mu][cb][volume], u_tmp is [cb][volume][mu];
// the matrices are copied whole (no transposition).
void pack_gauge_field(int volume, sse_u_mat_array* u, sse_u_mat_array* u_tmp) {
  for (int cb = 0; cb < 2; ++cb)
    for (int ix = 0; ix < volume; ++ix)
      for (int mu = 0; mu < 4; ++mu)
        memcpy(u_tmp + (mu + 4 * (ix + (size_t)volume * cb)), u + ix + (size_t)volume * (cb + 2 * mu),
               sizeof(sse_u_mat_array));
}

// lib/packer_nopad.c:47-69: despite its name the reference's unpacker reads
// and writes the same [mu][cb][ix] positions (two sites at a time through a
// stack buffer), so the caller's u ends up a copy of u_tmp.  The same bytes are
// produced here.  (For an odd volume the reference's pairwise loop also touches
// the entry one past each [mu][cb] block, which for the last block is past the
// arrays; with every global dimension even, volume = V/2 is a multiple of 8.)
void unpack_gauge_field(int volume, sse_u_mat_array* u_tmp, sse_u_mat_array* u) {
  if (volume <= 0 || u == u_tmp) return;
  memmove(u, u_tmp, (size_t)8 * (size_t)volume * sizeof(sse_u_mat_array));
}

// The true inverse of pack_gauge_field: [cb][volume][mu] back to [mu][cb][volume].
void unpack_gauge_field_inverse(int volume, sse_u_mat_array* u_tmp, sse_u_mat_array* u) {
  for (int cb = 0; cb < 2; ++cb)
    for (int ix = 0; ix < volume; ++ix)
      for (int mu = 0; mu < 4; ++mu)
        memcpy(u + ix + (size_t)volume * (cb + 2 * mu), u_tmp + (mu + 4 * (ix + (size_t)volume * cb)),
               sizeof(sse_u_mat_array));
}
