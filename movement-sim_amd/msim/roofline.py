"""Algorithmic bytes of the fused tick, per agent, itemised per kernel.

These are the bytes the algorithm has to move, not what a given kernel
happens to move (rocprofv3 FETCH_SIZE/WRITE_SIZE give the latter):

  advance   28 B  read pos 8 + target 8 + speed 4, write pos 8
                  (24 B when all agents share the shader's constant speed)
  scan       8 B  bucket-offset traffic per agent (counts in, starts out)
  scatter   12 B  binned position 8 + agent id 4 written
  collide   16 B  binned position 8 + agent id 4 read back, count 4 written
  ---------------
  fused     64 B  per agent-tick

The arrival path (road 32 B, connections <= 24 B, xorshift state 32 B, road
and target writes 12 B) runs for a few percent of agent-ticks and is reported
separately as arrivals per agent-tick.
"""

HBM_PEAK_BYTES_PER_S = 8.0e12  # MI355X HBM3E

B_ADVANCE = 28
B_ADVANCE_UNIFORM_SPEED = 24
B_SCAN = 8
B_SCATTER = 12
B_COLLIDE = 16
B_FUSED = B_ADVANCE + B_SCAN + B_SCATTER + B_COLLIDE
assert B_FUSED == 64



def kernel_bytes(name, uniform_speed=False):
    """Algorithmic bytes per agent for one launch of the named kernel.

    k_advance / k_advance4 28 (24 with the shader's constant speed); k_bin_count 8 (bucket
    offsets) and k_bin_place 12 (binned position + id) in the segment path, or
    k_scan_down 8 and k_scatter 12 in the atomic path, k_bin_fused 20 (both)
    in the one-pass segment path; k_collide / k_collide_staged 16.  The
    batched kernels k_bbin_count, k_bbin_place, k_bbin_fused, k_bcollide and
    k_bcollide_staged do the same work per agent as their plain
    counterparts.  k_move_sorted / k_bmove_sorted (advance and rebinning in
    one pass over the bucket-sorted state) 48 (44 with the constant speed).
    k_route (arrival path), the single-workgroup scan/finish passes,
    k_bfinish and k_reorder: 0."""
    name = name.split("<")[0]
    if name in ("k_move_sorted", "k_bmove_sorted"):
        return (B_ADVANCE_UNIFORM_SPEED if uniform_speed else B_ADVANCE) + B_SCAN + B_SCATTER
    if name in ("k_advance", "k_advance4", "k_advance_bin"):
        return B_ADVANCE_UNIFORM_SPEED if uniform_speed else B_ADVANCE
    if name in ("k_bin_count", "k_bbin_count", "k_scan_down"):
        return B_SCAN
    if name in ("k_bin_place", "k_bbin_place", "k_scatter"):
        return B_SCATTER
    if name in ("k_bin_fused", "k_bbin_fused"):
        return B_SCAN + B_SCATTER
    if name in ("k_collide", "k_bcollide", "k_collide_staged", "k_bcollide_staged"):
        return B_COLLIDE
    return 0
