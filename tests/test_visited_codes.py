"""The router's visited cache codes (router_device.hpp vis_key) identify a
vertex exactly: for every vid below 2^B, (slot, code) of choice 1 and of
choice 2 are each injective, and the two choices' code ranges do not overlap,
so a lookup can only hit on the vertex that was stored.  This is the property
that keeps search results independent of the cache (a false hit would skip a
vertex the reference visits)."""
import numpy as np
import pytest


def vis_bytes(K, u8):
    # router_kernels.hpp vis_bytes
    if u8:
        return 7680 if K <= 4 else (5632 if K <= 8 else 4224)
    return 10240 if K <= 4 else (8704 if K <= 8 else (6144 if K <= 13 else 4608))


def codes(v, B, mode, K, u8):
    # Same arithmetic as vis_key<K, U8> in artea_amd/csrc/router_device.hpp.
    s8, s16 = vis_bytes(K, u8), vis_bytes(K, u8) // 2
    c8, c16 = (2**20 - 1) // s8 + 1, (2**27 - 1) // s16 + 1
    hm = np.uint32((1 << B) - 1)
    h1 = (v * np.uint32(0x9E3779B1)) & hm
    h2 = (v * np.uint32(0x85EBCA77)) & hm
    if mode == 0:
        q1, q2 = h1 // s8, h2 // s8
        return h1 - q1 * s8, q1 + 1, h2 - q2 * s8, np.where(q2 + c8 + 1 <= 255, q2 + c8 + 1, 0)
    q1, q2 = h1 // s16, h2 // s16
    return h1 - q1 * s16, q1 + 1, h2 - q2 * s16, np.where((mode == 1) & (q2 + c16 + 1 <= 65535), q2 + c16 + 1, 0)


@pytest.mark.parametrize("u8", [True, False])
@pytest.mark.parametrize("K", [4, 8, 13, 16])
@pytest.mark.parametrize("B,mode", [(13, 0), (20, 0), (21, 1), (24, 1)])
def test_visited_codes_identify_the_vertex(B, mode, K, u8):
    v = np.arange(1 << B, dtype=np.uint32)
    s1, t1, s2, t2 = codes(v, B, mode, K, u8)
    width = 256 if mode == 0 else 65536
    slots = vis_bytes(K, u8) if mode == 0 else vis_bytes(K, u8) // 2
    assert s1.max() < slots and s2.max() < slots
    assert t1.min() >= 1 and t1.max() < width
    assert len(np.unique(s1.astype(np.uint64) * width + t1)) == len(v)
    second = t2 != 0
    if second.any():
        assert t2[second].max() < width and t2[second].min() > t1.max()
        assert len(np.unique(s2[second].astype(np.uint64) * width + t2[second])) == second.sum()
    if K <= 4:
        assert second.mean() > 0.85  # most vertices have a second choice with the largest caches


def test_visited_code_bounds_at_the_mode_limits():
    # The choice-1 codes of every cache size fit their width with room left.
    for K in (4, 8, 13, 16):
        for u8 in (True, False):
            assert (2**20 - 1) // vis_bytes(K, u8) + 1 < 255
            assert (2**27 - 1) // (vis_bytes(K, u8) // 2) + 1 < 65535
