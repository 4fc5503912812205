//----------------------------------*-C++-*----------------------------------//
// celeritas_amd: MI355X-native per-step transport loop
//---------------------------------------------------------------------------//
//! \file geo/GeoParams.hh
//---------------------------------------------------------------------------//
#pragma once

#include <string>
#include <vector>

#include "corecel/DeviceRuntime.hh"
#include "corecel/Json.hh"
#include "GeoData.hh"

namespace celeritas_amd
{
//---------------------------------------------------------------------------//
/*!
 * Host description of one simple unit (the ORANGE JSON "simple unit").
 */
struct GeoInput
{
    std::vector<std::uint8_t> surf_types;
    std::vector<std::vector<double>> surf_data;
    std::vector<std::string> surface_labels;

    struct Volume
    {
        std::vector<std::uint32_t> faces;
        std::vector<std::int32_t> logic;
        std::uint8_t flags{0};
        std::string label;
        bool background{false};
    };
    std::vector<Volume> volumes;
};

//! Parse the first universe of an ORANGE ".org.json" document
GeoInput geo_input_from_orange_json(JsonValue const& doc);

//! Convert an ORANGE postfix logic string ("0 1 ~ & 2 &") to tokens
std::vector<std::int32_t> parse_logic_string(std::string const& s);

//---------------------------------------------------------------------------//
/*!
 * Geometry params: validated host data plus device arrays.
 *
 * Replaces OrangeParams (orange/OrangeParams.hh) for one simple unit.
 * Per-volume bounding boxes come from the CSG logic (interval propagation
 * through the postfix expression); point location walks a bounding
 * interval hierarchy over them (the reference's BIH, orange/detail/
 * BIHBuilder.cc, stored depth-first with escape indices so the device
 * walk needs no stack) and tests a volume's box before its senses.
 */
class GeoParams
{
  public:
    explicit GeoParams(GeoInput const& inp);

    GeoParamsRef const& device_ref() const { return ref_; }
    size_type num_volumes() const { return num_volumes_; }
    size_type num_surfaces() const { return num_surfaces_; }
    std::string const& volume_label(size_type v) const { return labels_[v]; }
    //! Find a volume by label (invalid_id if absent)
    std::uint32_t find_volume(std::string const& label) const;

    // Host copies (also exported for the test harness)
    std::vector<std::uint8_t> h_surf_type;
    std::vector<std::uint32_t> h_surf_offset;
    std::vector<double> h_surf_data;
    std::vector<std::uint32_t> h_face_begin;
    std::vector<std::uint8_t> h_num_faces;
    std::vector<std::uint8_t> h_max_isect;
    std::vector<std::uint32_t> h_faces;
    std::vector<std::uint32_t> h_logic_begin;
    std::vector<std::uint16_t> h_logic_size;
    std::vector<std::int32_t> h_logic;
    std::vector<std::uint8_t> h_flags;
    std::vector<double> h_bbox;
    std::vector<std::uint32_t> h_conn_begin;
    std::vector<std::uint32_t> h_conn_size;
    std::vector<std::uint32_t> h_conn_vols;
    std::vector<BihNode> h_bih_nodes;
    std::vector<std::uint32_t> h_bih_vols;
    std::uint32_t background{invalid_id};

  private:
    std::uint32_t bih_inf_begin_{0};
    std::uint32_t bih_inf_count_{0};
    size_type num_volumes_{0};
    size_type num_surfaces_{0};
    std::vector<std::string> labels_;

    DeviceVector<char> d_hot_;  //!< the tables tracking kernels stage in LDS
    DeviceVector<BihNode> d_bih_nodes_;
    DeviceVector<std::uint32_t> d_bih_vols_;
    GeoParamsRef ref_;
};

//---------------------------------------------------------------------------//
}  // namespace celeritas_amd
