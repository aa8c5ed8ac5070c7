// save(fp, name, obj) / load(fp, name, obj) of the C++ API objects in MAS
// files, with the reference's block names so that checkpoints interchange
// with its tooling:
//   scalar            name                 (1 x 1)
//   std::string       name                 char (len x 1)
//   std::vector<T>    name                 (n x 1)            container_sl.hpp:13-30
//   std::vector<vec<D>> name               (n x D x 1)        container_sl.hpp:96-134
//   simple_grid       name.lower/.upper (double D x 1), name.cells (int32 D x 1)
//                                                             simple_grid.hpp:429-443
//   host_field        name.location (int32), name.grid.*, name.content
//                                                             host_field.hpp:310-337
//   device_field      name.location (int32), name.grid.*, name.content.device,
//                     name.content.content                   device_field.hpp:327-352
//   device_vector     name.device, name.content              device_vector.hpp:294-341
//   records           name.<tag_name> per tag (tag-wise SoA) tag_sl.hpp:62-137
//   particle_group    name.data (the stored records)         particle_group.hpp:272-283
#pragma once

#include <string>
#include <type_traits>
#include <vector>

#include "device_field.hpp"
#include "host_field.hpp"
#include "device_vector.hpp"
#include "grid.hpp"
#include "mas.hpp"
#include "particle_group.hpp"
#include "tag.hpp"

namespace psum {
namespace serialization {

namespace detail {
template <typename T>
struct is_vec : std::false_type {};
template <int D, typename S>
struct is_vec<psum::vec<D, S>> : std::true_type {};
template <typename T>
struct is_record : std::false_type {};
template <typename... B>
struct is_record<psum::tag::tagged_struct<B...>> : std::true_type {};

inline std::string device_label(hipStream_t)
{
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) return "[#0]";
    return std::string("[#") + std::to_string(dev) + "]" + prop.name;
}
}  // namespace detail

template <typename... B>
void save_records(mas_file& f, const std::string& name, const std::vector<psum::tag::tagged_struct<B...>>& v);
template <typename... B>
void load_records(mas_file& f, const std::string& name, std::vector<psum::tag::tagged_struct<B...>>& v);

template <typename T, typename = std::enable_if_t<std::is_arithmetic_v<T>>>
void save(mas_file& f, const std::string& name, T v)
{
    f.writeScalar<T>(name, v);
}
template <typename T, typename = std::enable_if_t<std::is_arithmetic_v<T>>>
void load(mas_file& f, const std::string& name, T& v)
{
    const std::vector<T> d = f.read<T>(name);
    if (d.size() != 1) throw std::runtime_error("MAS: block '" + name + "' is not a scalar");
    v = d[0];
}

inline void save(mas_file& f, const std::string& name, const std::string& s) { f.writeData(name, s); }
inline void load(mas_file& f, const std::string& name, std::string& s)
{
    const std::vector<char> d = f.read<char>(name);
    s.assign(d.begin(), d.end());
}

template <typename T>
void save(mas_file& f, const std::string& name, const std::vector<T>& v)
{
    if constexpr (std::is_arithmetic_v<T>) {
        f.writeData<T>(name, v.data(), {v.size(), 1});
    } else if constexpr (detail::is_vec<T>::value) {
        using S = std::remove_cv_t<std::remove_reference_t<decltype(v[0][0])>>;
        constexpr int D = T::size();
        std::vector<S> flat(v.size() * D);
        for (size_t i = 0; i < v.size(); ++i)
            for (int d = 0; d < D; ++d) flat[i * D + d] = v[i][d];
        f.writeData<S>(name, flat.data(), {v.size(), (size_t)D, 1});
    } else {
        static_assert(detail::is_record<T>::value, "MAS save: unsupported element type");
        save_records(f, name, v);
    }
}

template <typename T>
void load(mas_file& f, const std::string& name, std::vector<T>& v)
{
    if constexpr (std::is_arithmetic_v<T>) {
        std::vector<size_t> shp;
        v = f.read<T>(name, &shp);
        if (shp.size() != 2 || shp[1] != 1) throw std::runtime_error("MAS: block '" + name + "' is not a vector");
    } else if constexpr (detail::is_vec<T>::value) {
        using S = std::remove_cv_t<std::remove_reference_t<decltype(v.data()[0][0])>>;
        constexpr int D = T::size();
        std::vector<size_t> shp;
        const std::vector<S> flat = f.read<S>(name, &shp);
        if (shp.size() != 3 || shp[1] != (size_t)D || shp[2] != 1)
            throw std::runtime_error("MAS: block '" + name + "' is not a vector of " + std::to_string(D) + "-vectors");
        v.resize(shp[0]);
        for (size_t i = 0; i < v.size(); ++i)
            for (int d = 0; d < D; ++d) v[i][d] = flat[i * D + d];
    } else {
        static_assert(detail::is_record<T>::value, "MAS load: unsupported element type");
        load_records(f, name, v);
    }
}

// records, one block per tag
template <typename... B>
void save_records(mas_file& f, const std::string& name, const std::vector<psum::tag::tagged_struct<B...>>& v)
{
    auto one = [&](auto bind) {
        using Bind = decltype(bind);
        std::vector<typename Bind::type> comp(v.size());
        for (size_t i = 0; i < v.size(); ++i) comp[i] = psum::tag::get<typename Bind::tag>(v[i]);
        save(f, name + "." + Bind::tag::tag_name, comp);
    };
    (one(B{}), ...);
}
template <typename... B>
void load_records(mas_file& f, const std::string& name, std::vector<psum::tag::tagged_struct<B...>>& v)
{
    size_t n = (size_t)-1;
    auto one = [&](auto bind) {
        using Bind = decltype(bind);
        std::vector<typename Bind::type> comp;
        load(f, name + "." + Bind::tag::tag_name, comp);
        if (n == (size_t)-1) {
            n = comp.size();
            v.assign(n, {});
        } else if (comp.size() != n) {
            throw std::runtime_error("MAS: tags of '" + name + "' differ in length");
        }
        for (size_t i = 0; i < n; ++i) psum::tag::get<typename Bind::tag>(v[i]) = comp[i];
    };
    (one(B{}), ...);
}

template <int D>
void save(mas_file& f, const std::string& name, const field::simple_grid<D>& g)
{
    const psum_grid p = g.to_psum();
    std::vector<double> lo(p.lb, p.lb + D), up(p.ub, p.ub + D);
    std::vector<int32_t> cells(p.cells, p.cells + D);
    save(f, name + ".lower", lo);
    save(f, name + ".upper", up);
    save(f, name + ".cells", cells);
}
template <int D>
void load(mas_file& f, const std::string& name, field::simple_grid<D>& g)
{
    std::vector<double> lo, up;
    std::vector<int32_t> cells;
    load(f, name + ".lower", lo);
    load(f, name + ".upper", up);
    load(f, name + ".cells", cells);
    g = field::simple_grid<D>(lo, up, std::vector<int>(cells.begin(), cells.end()));
}

namespace detail {
// field content: scalars as an n x 1 block; rows of QS components as an
// n x 1 x QS block, the shape of a vector of Eigen::RowVector
// (container_sl.hpp:94-134)
template <typename Scalar, int QS>
void save_content(mas_file& f, const std::string& name, const std::vector<field::field_value_t<Scalar, QS>>& v)
{
    if constexpr (QS == 1) {
        save(f, name, v);
    } else {
        std::vector<Scalar> flat(v.size() * QS);
        for (size_t i = 0; i < v.size(); ++i)
            for (int q = 0; q < QS; ++q) flat[i * QS + q] = v[i][q];
        f.writeData<Scalar>(name, flat.data(), {v.size(), (size_t)1, (size_t)QS});
    }
}
template <typename Scalar, int QS>
void load_content(mas_file& f, const std::string& name, std::vector<field::field_value_t<Scalar, QS>>& v)
{
    if constexpr (QS == 1) {
        load(f, name, v);
    } else {
        std::vector<size_t> shp;
        const std::vector<Scalar> flat = f.read<Scalar>(name, &shp);
        if (shp.size() != 3 || shp[1] != 1 || shp[2] != (size_t)QS)
            throw std::runtime_error("MAS: block '" + name + "' is not a vector of " + std::to_string(QS) + "-rows");
        v.assign(shp[0], {});
        for (size_t i = 0; i < v.size(); ++i)
            for (int q = 0; q < QS; ++q) v[i][q] = flat[i * QS + q];
    }
}
}  // namespace detail

template <int D, field::var_loc Loc, typename Scalar, int QS>
void save(mas_file& f, const std::string& name, const field::device_field<D, Loc, Scalar, QS>& fld)
{
    save(f, name + ".location", (int32_t)Loc);
    save(f, name + ".grid", fld.getGrid());
    save(f, name + ".content.device", detail::device_label(nullptr));
    detail::save_content<Scalar, QS>(f, name + ".content.content", fld.to_host());
}
// the field keeps its grid: the stored grid must have the same cells
template <int D, field::var_loc Loc, typename Scalar, int QS>
void load(mas_file& f, const std::string& name, field::device_field<D, Loc, Scalar, QS>& fld)
{
    int32_t loc = -1;
    load(f, name + ".location", loc);
    if (loc != (int32_t)Loc) throw std::runtime_error("MAS: block '" + name + "' has another location");
    field::simple_grid<D> g;
    load(f, name + ".grid", g);
    for (int d = 0; d < D; ++d)
        if (g.to_psum().cells[d] != fld.getGrid().to_psum().cells[d])
            throw std::runtime_error("MAS: block '" + name + "' has another grid");
    std::vector<field::field_value_t<Scalar, QS>> c;
    detail::load_content<Scalar, QS>(f, name + ".content.content", c);
    if (c.size() != fld.size()) throw std::runtime_error("MAS: block '" + name + "' content does not match its grid");
    fld.copy(c);
}

// host_field: name.location, name.grid.*, name.content (host_field.hpp:310-337)
template <int D, field::var_loc Loc, typename Scalar, int QS>
void save(mas_file& f, const std::string& name, const field::host_field<D, Loc, Scalar, QS>& fld)
{
    save(f, name + ".location", (int32_t)Loc);
    save(f, name + ".grid", fld.getGrid());
    detail::save_content<Scalar, QS>(f, name + ".content", fld.getContent());
}
// the field takes the stored grid
template <int D, field::var_loc Loc, typename Scalar, int QS>
void load(mas_file& f, const std::string& name, field::host_field<D, Loc, Scalar, QS>& fld)
{
    int32_t loc = -1;
    load(f, name + ".location", loc);
    if (loc != (int32_t)Loc) throw std::runtime_error("MAS: block '" + name + "' has another location");
    field::simple_grid<D> g;
    load(f, name + ".grid", g);
    std::vector<field::field_value_t<Scalar, QS>> c;
    detail::load_content<Scalar, QS>(f, name + ".content", c);
    if (c.size() != g.contentSize(Loc)) throw std::runtime_error("MAS: block '" + name + "' content does not match its grid");
    fld.reset(g, std::move(c));
}

template <typename T>
void save(mas_file& f, const std::string& name, const particle_container::device_vector<T>& dv)
{
    save(f, name + ".device", detail::device_label(dv.get_queue()));
    save(f, name + ".content", dv.to_host());
}
template <typename T>
void load(mas_file& f, const std::string& name, particle_container::device_vector<T>& dv)
{
    std::vector<T> h;
    load(f, name + ".content", h);
    dv = particle_container::device_vector<T>(dv.get_queue(), h);
}

template <typename P, template <typename> typename V>
void save(mas_file& f, const std::string& name, const particle_container::particle_group<P, V>& g)
{
    save(f, name + ".data", g.get_content());
}
template <typename P, template <typename> typename V>
void load(mas_file& f, const std::string& name, particle_container::particle_group<P, V>& g)
{
    std::vector<P> h;
    load(f, name + ".data.content", h);
    g.clear();
    g.insert(h);
}

}  // namespace serialization
}  // namespace psum
