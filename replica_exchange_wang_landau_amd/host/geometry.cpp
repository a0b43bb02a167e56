#include "geometry.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <filesystem>
#include <fstream>
#include <sstream>
#include <stdexcept>

namespace rewl::host {
namespace {

double norm(Vec3 a) { return std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }

Vec3 unit(Vec3 a)
{
    const double length = norm(a);
    if (!(length > 0.0) || !std::isfinite(length))
        throw std::invalid_argument("Ising axes must be finite non-zero vectors");
    return {a.x / length, a.y / length, a.z / length};
}

std::string trim_field(const std::string &value)
{
    const auto first = value.find_first_not_of(" \t\r\n");
    if (first == std::string::npos) return {};
    const auto last = value.find_last_not_of(" \t\r\n");
    std::string s = value.substr(first, last - first + 1);
    // UTF-8 byte-order mark at the start of a field
    if (s.size() >= 3 && static_cast<unsigned char>(s[0]) == 0xef &&
        static_cast<unsigned char>(s[1]) == 0xbb && static_cast<unsigned char>(s[2]) == 0xbf)
        s.erase(0, 3);
    return s;
}

} // namespace

void Geometry::validate() const
{
    if (positions.empty()) throw std::invalid_argument("Geometry contains no sites");
    if (positions.size() != axes.size())
        throw std::invalid_argument("Geometry positions/axes size mismatch");
    if (periodic) {
        for (double l : {box.x, box.y, box.z})
            if (!std::isfinite(l) || !(l >= 0.0))
                throw std::invalid_argument("Periodic box lengths must be finite and non-negative");
        if (box.x == 0.0 && box.y == 0.0 && box.z == 0.0)
            throw std::invalid_argument("At least one periodic box length must be positive");
    }
    for (std::size_t i = 0; i < positions.size(); ++i) {
        const Vec3 &p = positions[i];
        if (!std::isfinite(p.x) || !std::isfinite(p.y) || !std::isfinite(p.z))
            throw std::invalid_argument("Site coordinates must be finite");
        if (std::abs(norm(axes[i]) - 1.0) > 1e-10)
            throw std::invalid_argument("Ising axes must be normalized");
    }
}

Geometry Geometry::simple_cubic(std::size_t nx, std::size_t ny, std::size_t nz, double spacing,
                                Vec3 axis, bool periodic)
{
    if (nx == 0 || ny == 0 || nz == 0 || !(spacing > 0.0))
        throw std::invalid_argument("Invalid simple-cubic dimensions or spacing");
    Geometry g;
    g.box = {static_cast<double>(nx) * spacing, static_cast<double>(ny) * spacing,
             static_cast<double>(nz) * spacing};
    g.periodic = periodic;
    const Vec3 e = unit(axis);
    g.positions.reserve(nx * ny * nz);
    for (std::size_t z = 0; z < nz; ++z)
        for (std::size_t y = 0; y < ny; ++y)
            for (std::size_t x = 0; x < nx; ++x)
                g.positions.push_back({static_cast<double>(x) * spacing,
                                       static_cast<double>(y) * spacing,
                                       static_cast<double>(z) * spacing});
    g.axes.assign(g.positions.size(), e);
    g.validate();
    return g;
}

Geometry Geometry::load_xyz_axes(const std::string &path, Vec3 box, bool periodic)
{
    std::ifstream input(path);
    if (!input) throw std::runtime_error("Cannot open geometry file: " + path);
    Geometry g;
    g.box = box;
    g.periodic = periodic;
    std::string line;
    std::size_t number = 0;
    while (std::getline(input, line)) {
        ++number;
        const auto first = line.find_first_not_of(" \t\r");
        if (first == std::string::npos || line[first] == '#') continue;
        std::istringstream values(line);
        Vec3 p, a;
        if (!(values >> p.x >> p.y >> p.z >> a.x >> a.y >> a.z))
            throw std::runtime_error("Invalid geometry line " + std::to_string(number));
        g.positions.push_back(p);
        g.axes.push_back(unit(a));
    }
    g.validate();
    return g;
}

Geometry Geometry::load_csv(const std::string &path, Vec3 box, bool periodic)
{
    std::ifstream input(path);
    if (!input) throw std::runtime_error("Cannot open CSV geometry file: " + path);
    Geometry g;
    g.box = box;
    g.periodic = periodic;
    static const std::array<const char *, 6> header{"x", "y", "z", "mx", "my", "mz"};
    bool have_header = false;
    std::string line;
    std::size_t number = 0;
    while (std::getline(input, line)) {
        ++number;
        const std::string clean = trim_field(line);
        if (clean.empty() || clean[0] == '#') continue;
        std::vector<std::string> cells;
        {
            std::stringstream stream(clean);
            std::string cell;
            while (std::getline(stream, cell, ',')) cells.push_back(trim_field(cell));
        }
        if (!have_header) {
            bool ok = cells.size() == header.size();
            for (std::size_t k = 0; ok && k < header.size(); ++k) {
                std::string name = cells[k];
                std::transform(name.begin(), name.end(), name.begin(),
                               [](unsigned char c) { return static_cast<char>(std::tolower(c)); });
                ok = name == header[k];
            }
            if (!ok) throw std::runtime_error("CSV geometry header must be x,y,z,mx,my,mz");
            have_header = true;
            continue;
        }
        if (cells.size() != 6)
            throw std::runtime_error("CSV geometry line " + std::to_string(number) +
                                     " must have 6 columns");
        double v[6];
        for (std::size_t k = 0; k < 6; ++k) {
            std::size_t used = 0;
            const std::string message = "Invalid number at CSV geometry line " + std::to_string(number);
            try {
                v[k] = std::stod(cells[k], &used);
            } catch (const std::exception &) {
                throw std::runtime_error(message);
            }
            if (used != cells[k].size() || !std::isfinite(v[k])) throw std::runtime_error(message);
        }
        g.positions.push_back({v[0], v[1], v[2]});
        g.axes.push_back({v[3], v[4], v[5]});
    }
    if (!have_header) throw std::runtime_error("CSV geometry file has no header");
    g.validate();
    return g;
}

std::vector<double> Geometry::flat_positions() const
{
    std::vector<double> out;
    out.reserve(3 * positions.size());
    for (const auto &p : positions) out.insert(out.end(), {p.x, p.y, p.z});
    return out;
}

std::vector<double> Geometry::flat_axes() const
{
    std::vector<double> out;
    out.reserve(3 * axes.size());
    for (const auto &a : axes) out.insert(out.end(), {a.x, a.y, a.z});
    return out;
}

Geometry geometry_from_config(const RunConfig &c)
{
    if (c.geometry_file.empty())
        return Geometry::simple_cubic(c.nx, c.ny, c.nz, c.spacing, c.axis, c.periodic);
    if (std::filesystem::path(c.geometry_file).extension() == ".csv")
        return Geometry::load_csv(c.geometry_file, c.box_lengths, c.periodic);
    return Geometry::load_xyz_axes(c.geometry_file, c.box_lengths, c.periodic);
}

} // namespace rewl::host
