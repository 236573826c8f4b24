"""Host side of the final light: the C structs' layout as the binding sees
it, the room grid's final-light inputs and face neighbours."""
from __future__ import annotations

import ctypes
import os
import shutil
import subprocess

import numpy as np
import pytest

from final_cases import oracle_faces, random_case
from hl2sdk_tools_sharp_amd import rooms
from hl2sdk_tools_sharp_amd.final import FACE_DTYPE, VhFinalDesc, VhFinalOpts

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PROBE = r"""
#include <stddef.h>
#include <stdio.h>
#include "vradhip.h"
int main(void) {
    printf("%zu %zu %zu %zu %zu %zu\n", sizeof(vh_final_face), offsetof(vh_final_face, min_s),
           offsetof(vh_final_face, minlight), offsetof(vh_final_face, lit), sizeof(vh_final_desc),
           offsetof(vh_final_desc, npatch));
    printf("%zu %zu\n", sizeof(vh_final_opts), offsetof(vh_final_opts, red_errors));
    return 0;
}
"""


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_struct_layouts_match_the_header(tmp_path):
    src = tmp_path / "probe.c"
    src.write_text(PROBE)
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    a, b = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")[:2]
    size, o_min, o_ml, o_lit, dsize, o_np = map(int, a.split())
    assert size == FACE_DTYPE.itemsize
    assert (o_min, o_ml, o_lit) == (FACE_DTYPE.fields["min_s"][1], FACE_DTYPE.fields["minlight"][1],
                                     FACE_DTYPE.fields["lit"][1])
    assert dsize == ctypes.sizeof(VhFinalDesc) and o_np == VhFinalDesc.npatch.offset
    osize, o_red = map(int, b.split())
    assert osize == ctypes.sizeof(VhFinalOpts) and o_red == VhFinalOpts.red_errors.offset


@pytest.fixture(scope="module")
def grid():
    g = rooms.room_grid(2, 2, tile=64.0, seed=2)
    return g, rooms.room_faces(g, 16.0, bumped_every=5)


def test_room_samples_sit_on_their_own_luxels(grid):
    g, fset = grid
    ff, fp = rooms.room_final_inputs(g, fset)
    faces = oracle_faces(ff, fset)
    n = ff.faces["width"].astype(np.int64) * ff.faces["height"]
    assert ff.luxel_pos is None
    assert np.array_equal(np.diff(fset.face_sample_ptr), np.where(ff.faces["lit"] == 1, n, 0))
    for f in range(0, fset.nfaces, 13):
        face = faces[f]
        for k in range(fset.face_sample_ptr[f], fset.face_sample_ptr[f + 1]):
            s, t = face.world_to_luxel(fset.sample_pos[k])
            j = k - fset.face_sample_ptr[f]
            assert (int(s + np.float32(0.5)), int(t + np.float32(0.5))) == (j % face.w, j // face.w)
            b = ff.sample_bounds[k]
            assert b[0] < s < b[2] and b[1] < t < b[3]
            # the grid point's luxel (BuildFaceLuxels) is within rounding of the sample
            lw = face.luxel_to_world(j % face.w, j // face.w)
            assert np.allclose(lw, fset.sample_pos[k], atol=1e-3)
    assert fp.npatch == g.npatch and fp.winding_ptr[-1] == 4 * g.npatch


def test_face_neighbours_share_a_corner_and_a_plane(grid):
    g, _ = grid
    ptr, nbr = rooms.face_neighbours(g)
    pairs = {(f, int(x)) for f in range(g.nface) for x in nbr[ptr[f]:ptr[f + 1]]}
    assert all((b, a) in pairs for a, b in pairs)
    assert all(a != b for a, b in pairs)
    for a, b in list(pairs)[::17]:
        assert np.array_equal(g.face_normal[a], g.face_normal[b])
        ca = [(x, y, z) for x in (g.face_lo[a][0], g.face_hi[a][0]) for y in (g.face_lo[a][1], g.face_hi[a][1])
              for z in (g.face_lo[a][2], g.face_hi[a][2])]
        cb = [(x, y, z) for x in (g.face_lo[b][0], g.face_hi[b][0]) for y in (g.face_lo[b][1], g.face_hi[b][1])
              for z in (g.face_lo[b][2], g.face_hi[b][2])]
        assert set(ca) & set(cb)
    # a floor tile in the middle of a room has its eight tiles round it
    assert np.max(np.diff(ptr)) == 8
    for f in range(g.nface):
        assert list(nbr[ptr[f]:ptr[f + 1]]) == sorted(nbr[ptr[f]:ptr[f + 1]])


def test_random_case_is_well_formed():
    casters, L, fset, ff, fp = random_case(1)
    n = ff.faces["width"].astype(np.int64) * ff.faces["height"]
    assert len(ff.luxel_pos) == n.sum()
    assert len(ff.sample_bounds) == fset.nsamples
    assert ff.nbr_ptr[-1] == len(ff.nbr) and np.all((ff.nbr >= 0) & (ff.nbr < fset.nfaces))
    assert fp.winding_ptr[-1] == len(fp.winding)
    assert np.any(fp.child1 >= 0) and np.any(fset.face_normal_count == 4)
    assert np.any(ff.faces["lit"] == 0) and np.any(np.diff(fset.face_sample_ptr) == 0)
