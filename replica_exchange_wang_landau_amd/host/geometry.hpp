// Site geometry of the dipole system (Geometry, physics.hpp:12-35): positions, unit Ising
// axes and an orthorhombic box whose zero-length axes are open.
#pragma once

#include "config.hpp"

#include <string>
#include <vector>

namespace rewl::host {

struct Geometry {
    std::vector<Vec3> positions;
    std::vector<Vec3> axes;
    Vec3 box{};
    bool periodic{true};

    std::size_t size() const { return positions.size(); }
    void validate() const; // physics.cpp:34-56
    // site index x + nx*(y + ny*z), box = n*spacing (physics.cpp:58-75)
    static Geometry simple_cubic(std::size_t nx, std::size_t ny, std::size_t nz, double spacing,
                                 Vec3 axis, bool periodic);
    // whitespace "x y z mx my mz" lines, '#' comments, axes normalised (physics.cpp:77-97)
    static Geometry load_xyz_axes(const std::string &path, Vec3 box, bool periodic);
    // header x,y,z,mx,my,mz, axes validated not normalised (physics.cpp:99-149)
    static Geometry load_csv(const std::string &path, Vec3 box, bool periodic);
    // row-major N x 3 copies for the C ABI
    std::vector<double> flat_positions() const;
    std::vector<double> flat_axes() const;
};

// the geometry wl_run builds from its configuration (wl_run.cpp:45-51)
Geometry geometry_from_config(const RunConfig &config);

} // namespace rewl::host
