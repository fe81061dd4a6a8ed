// Host preparation of the device layout of one solver handle (plain C++,
// no HIP types): mesh validation, the Hilbert relabelling, the SoA arrays of
// both paths and the tiles of the tiled step.  Built with g++ and OpenMP and
// linked into librhllc_hip.so.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../mesh/mesh_common.h"

namespace rhllc_layout {

struct Options {
  bool mixed = false;           // face vectors S n and 1/V instead of (n, S) and V
  bool reorder = true;          // Hilbert relabelling of cells (faces are grouped either way when tiled)
  bool tiled = true;            // build the tiles of the tiled step (else the face index per cell slot)
  bool has_out_source = false;  // boundary code -3 faces are allowed (fixed ghost state)
  bool has_inner_source = false;  // code -4 takes the second fixed ghost state (else copies)
  bool lock_wall = false;         // code -2 mirrors the normal velocity (else copies)
  bool reject_outer_surface = false;  // code -5 faces are an error
  bool out_source_copy = false;       // code -3 copies the cell when there is no out_source state
  int tile_cells = 256;
};

struct Layout {
  std::vector<int32_t> cell_new_to_old, face_new_to_old;
  // per face (left, right): right = the neighbour; N and N + 1 the ghost
  // states of codes -3 and -4; ~left (negative) a wall of code -2; left for
  // the copy-type codes
  rhllc_mesh::RawVec<int32_t> lr;
  rhllc_mesh::RawVec<double> nS;       // per face (nx, ny, nz, S), FP64 path
  // per face S n in 16 bytes, mixed path: the FP32 vector and, in the fourth
  // word, three signed 7-bit multiples (bits 0-6, 7-13, 14-20) of
  // Q = 2^(E - 30), 2^E the power of two at or below the largest |component|
  // of the FP32 vector, that restore what the FP32 rounding took: fp32 + k Q
  // is the vector to 2^-30 of its largest component, so a cell's four vectors
  // close to that (FP32 alone: 2^-24).  On the tiled step bits 21-28 hold the
  // entry's cell in its tile (cell - 256 t) and bit 29 whether that is the
  // right cell; far holds the other cell (or the right side's boundary
  // encoding), so (left, right) = bit 29 ? (far, near) : (near, far).
  rhllc_mesh::RawVec<float> Sn;
  rhllc_mesh::RawVec<int32_t> far;     // tiled mixed path, F + extra entries
  rhllc_mesh::RawVec<double> V;        // per cell, FP64 path
  rhllc_mesh::RawVec<double> Vinv;     // per cell 1/V, mixed path
  rhllc_mesh::RawVec<int32_t> cf;      // per cell 4 face indices | bit 31 when the cell is id_r (two-kernel step)
  // tiled step
  int32_t n_tiles = 0, max_slots = 0;
  std::vector<int32_t> tile_meta;    // per tile (first own face, own count, first extra, extra count)
  // extra faces of all tiles, tile by tile, by face index.  lr and nS / Sn
  // continue past the F faces with a copy for each entry, in this order
  // (entry k at F + k), so a tile reads its extra faces contiguously and
  // through the same arrays as its own.
  std::vector<int32_t> tile_extra;
  rhllc_mesh::RawVec<uint32_t> cell_slots;  // per cell two words: four 16-bit slots, bit 15 = id_r
};

// One face vector S n into the mixed path's 16-byte form (Layout::Sn).
void pack_sn(const double e[3], float *o);

// Validates the reference's geo_cell_t / geo_face_t records (every id in
// range, boundary codes known, each cell slot agreeing with its face's
// id_l / id_r and sign bit, V > 0) and builds the layout.  Returns false with
// a message naming the first offending record.
bool build(const Options &opt, int64_t n_cells, const void *geo_cells, int64_t n_faces, const void *geo_faces,
           Layout &out, std::string &err);

}  // namespace rhllc_layout
