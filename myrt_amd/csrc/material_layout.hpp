// Named-parameter material layouts: the reference's object_type / types_registry
// (src/types.hpp:216-310) and material_registry / material_buffer (src/material.hpp:10-118).
//
// A layout is a list of named parameters, each a built-in constant type (float ... mat4,
// vec4unorm8) or an earlier layout; parameters are packed as consecutive 4-byte blocks in
// declaration order, nested layouts inline. Every material owns `blocks()` blocks of one flat
// buffer at its offset, zero-initialised, and set_parameter writes a parameter's object
// representation there. The shaders do not read that buffer: at upload every material is
// resolved through its material type's bindings into the 32-B record the GGX model reads
// (albedo, ior, roughness, metallic, transmission, albedo texture: the inputs of
// src/pt/material_new.glsl:17-25 `open()`).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "pathtrace.hpp"

namespace mrt {

// Built-in constant types in the reference's parameter_type_id order (src/types.hpp:27-44).
enum ConstKind : int32_t {
  kFloat = 0, kInt, kUint, kVec2, kVec3, kVec4, kVec4Unorm8, kIVec2, kIVec3, kIVec4,
  kUVec2, kUVec3, kUVec4, kMat2, kMat3, kMat4, kConstKinds
};
const char* const_kind_name(int32_t k);
uint32_t const_kind_blocks(int32_t k);

// A parameter's type: constant >= 0 names a ConstKind, otherwise `object` is a layout id.
struct TypeRef {
  int32_t constant = -1;
  uint32_t object = 0;
};

struct LayoutParam {
  std::string name;
  uint32_t offset = 0;  // in blocks from the start of the enclosing layout
  uint32_t blocks = 0;
  TypeRef type;
};

struct Layout {
  std::string name;
  uint32_t blocks = 0;
  std::vector<LayoutParam> params;
  bool sealed = false;  // used by a material type or another layout: no more parameters
  bool hidden = false;  // the built-in layouts behind mrt_material, not found by name
};

// Inputs of the GGX shading model.
enum ShadingInput { kInAlbedo = 0, kInIor, kInRoughness, kInMetallic, kInTransmission, kInTexture, kInputs };
const char* shading_input_name(int i);

struct InputBinding {
  int32_t offset = -1;  // block offset within the layout, -1 = unbound (the input's default)
  int32_t kind = -1;    // ConstKind of the bound block(s)
};

struct MaterialType {
  uint32_t layout = 0;
  InputBinding in[kInputs];
};

struct MaterialSlot {
  uint32_t type = 0;    // material type id
  uint32_t offset = 0;  // first block in MaterialStore::blocks
};

class MaterialStore {
 public:
  MaterialStore();

  int add_layout(const std::string& name, uint32_t* out);
  int add_parameter(uint32_t layout, const std::string& name, const std::string& type_name);
  // types_registry::find: the first layout of that name, else the constant type of that
  // identifier.
  bool find_type(const std::string& name, TypeRef* out) const;
  // paths[i] is the parameter path ("albedo.value") feeding input i, or null for the default
  // binding by name; an empty string leaves the input unbound.
  int push_material_type(uint32_t layout, const char* const* paths, uint32_t* out);
  int push_material(uint32_t type, uint32_t* out);
  // material_buffer::set: a parameter not in the layout is ignored (*found = false); a size other
  // than blocks * 4 bytes is an error.
  int set_raw(uint32_t material, const std::string& name, const void* data, size_t bytes, bool* found);
  // Float convenience: vec4unorm8 takes 3 or 4 values in [0, 1], integer kinds take integral
  // values, float kinds `blocks` values.
  int set_floats(uint32_t material, const std::string& name, const float* v, size_t n, uint32_t num_textures);
  // The whole record through the bindings; an unbound input must keep its default.
  int set_record(uint32_t material, const Material& m);
  Material resolve(uint32_t material) const;

  std::vector<Layout> layouts;
  std::vector<MaterialType> types;
  std::vector<MaterialSlot> materials;
  std::vector<uint32_t> blocks;  // material_buffer::buffer(), as raw 4-byte blocks

  static constexpr uint32_t kRecordType = 0;  // material type of mrt_material records

 private:
  const LayoutParam* find_param(uint32_t layout, const std::string& name) const;
  int resolve_path(uint32_t layout, const std::string& path, InputBinding* out) const;
  int32_t leaf_kind(const LayoutParam& p, uint32_t* offset) const;
};

}  // namespace mrt
