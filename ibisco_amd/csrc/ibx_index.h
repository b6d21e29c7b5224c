// Indices of the loads the kernels issue unconditionally.
//
// A load under a per-lane condition makes the compiler wait for it before it
// issues the next, so the per-step kernels load first and test afterwards:
// where the condition fails, the index falls back to element 0 and the value
// is dropped.  Element 0 exists in every array a kernel reads this way, because
// an empty host array is uploaded as one zeroed element (upload_count).  These
// functions are the only place the fallbacks are written; tests/index_model.cpp
// runs them over edge-case layouts on the host under AddressSanitizer.
#pragma once

namespace ibx {

// Bonded term meta word: kind (2 bits) | table << 2 | virtual-site members << kTermVsShift (-1 for padding).
constexpr int kTermVsShift = 26;
constexpr int kTermTabMask = (1 << (kTermVsShift - 2)) - 1;

// Elements a device array holds for n host elements.
__host__ __device__ inline long upload_count(long n) { return n > 0 ? n : 1; }

// List build: entry of exclusion slot e of an item whose exclusions are excl[ex0, ex1).
__host__ __device__ inline int excl_entry(int ex0, int e, int ex1) { return ex0 + e < ex1 ? ex0 + e : 0; }
// ... and the item whose sorted slot it looks up (item 0 past the item's own exclusions).
__host__ __device__ inline int excl_item(int ex0, int e, int ex1, int item) { return ex0 + e < ex1 ? item : 0; }

// Integrator: contribution slot q of a range ending at b.
__host__ __device__ inline int contrib_entry(int q, int b) { return q < b ? q : 0; }

// Bonded terms: item read for member m (-1 unused) of a term (meta -1: padding).
__host__ __device__ inline int term_member(int meta, int m) { return (meta >= 0 && m >= 0) ? m : 0; }
// ... and the term's table.
__host__ __device__ inline int term_table(int meta) { return ((meta > 0 ? meta : 0) >> 2) & kTermTabMask; }

// Integrator: the list slot of an item in a group it may not belong to (-1).
__host__ __device__ inline int member_slot(int slot) { return slot > 0 ? slot : 0; }

}  // namespace ibx
