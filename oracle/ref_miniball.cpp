// Test infrastructure only: a C entry point over the reference's own
// Miniball (Gaertner's header, vendored in the reference at
// projects/CUDA-Engine/source/Miniball.hpp, included from there by the
// Makefile's ref target), used the way BoundingSphere.cpp:10-48 uses it:
// float coordinates, a list of 3-vectors, the centre and squared radius.
// Built into oracle/_ref/ only, to check rsh_object_spheres against it.
#include <list>
#include <vector>

#include "Miniball.hpp"

extern "C" void ref_miniball(const float* pts, int n, float* center, float* squared_radius) {
    typedef float mytype;
    std::list<std::vector<mytype> > point_list;
    for (int i = 0; i < n; i++) point_list.push_back(std::vector<mytype>(pts + 3 * i, pts + 3 * i + 3));
    typedef std::list<std::vector<mytype> >::const_iterator PointIterator;
    typedef std::vector<mytype>::const_iterator CoordinateIterator;
    typedef Miniball::Miniball<Miniball::CoordAccessor<PointIterator, CoordinateIterator> > MB;
    MB mb(3, point_list.begin(), point_list.end());
    for (int j = 0; j < 3; j++) center[j] = mb.center()[j];
    *squared_radius = mb.squared_radius();
}
