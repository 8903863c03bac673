// Device octree (elfmm_octree.hip): the tree and lists of build_tree
// (elfmm_tree.cpp), built on the GPU from device coordinates.
#pragma once
#include <hip/hip_runtime.h>

#include "elfmm_internal.h"

namespace elfmm {

// Device buffers of the build, kept between builds (one per host thread).
struct OctreeWork {
    struct Impl;
    Impl *impl;
    OctreeWork();
    ~OctreeWork();
    OctreeWork(const OctreeWork &) = delete;
    OctreeWork &operator=(const OctreeWork &) = delete;
};

// Builds T from n sources and nt targets (device arrays, 3 doubles per
// point) on stream s; T is the same as build_tree's on the same points.
// too_deep is set (and T left unset) when a box below level 21 would split:
// the caller then builds the tree on the host.
hipError_t build_tree_dev(const double *d_src, int n, const double *d_tgt, int nt, int nbox, int imindlin,
                          Tree &T, OctreeWork &work, hipStream_t s, bool &too_deep);

}  // namespace elfmm
