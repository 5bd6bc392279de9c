// Named-parameter material layouts and their C ABI: object_type / types_registry
// (src/types.hpp:216-310), material_registry::push_type and material_buffer
// (src/material.hpp:18-118), scene::push_material_type / push_material / set_parameter
// (src/pathtracer/scene.hpp:131-140, 156, 249-252).
#include "material_layout.hpp"

#include <cmath>
#include <cstring>

#include "abi_guard.hpp"
#include "myrt.h"
#include "scene_internal.hpp"

namespace mrt {
namespace {

struct ConstInfo {
  const char* name;
  uint32_t blocks;
};
// Identifiers and block counts of the built-in types (src/types.hpp:48-98): one 4-byte block per
// component, vec4unorm8 one block of packed bytes, matrices column by column.
constexpr ConstInfo kConst[kConstKinds] = {
    {"float", 1}, {"int", 1},   {"uint", 1},  {"vec2", 2},  {"vec3", 3},  {"vec4", 4},
    {"vec4unorm8", 1}, {"ivec2", 2}, {"ivec3", 3}, {"ivec4", 4}, {"uvec2", 2}, {"uvec3", 3},
    {"uvec4", 4}, {"mat2", 4},  {"mat3", 9},  {"mat4", 16}};

constexpr const char* kInputNames[kInputs] = {"albedo", "ior", "roughness", "metallic", "transmission",
                                              "albedo_texture"};
constexpr uint32_t kMaxLayoutBlocks = 1u << 16;
constexpr size_t kMaxBlocks = size_t(1) << 30;

bool is_integer_kind(int32_t k) {
  return k == kInt || k == kUint || (k >= kIVec2 && k <= kUVec4);
}
bool is_unsigned_kind(int32_t k) { return k == kUint || (k >= kUVec2 && k <= kUVec4); }

bool input_accepts(int input, int32_t kind) {
  switch (input) {
    case kInAlbedo:
      return kind == kVec4Unorm8;  // the record holds albedo as RGBA8, like the reference's color_t
    case kInTexture:
      return kind == kUint || kind == kUVec2;
    default:
      return kind == kFloat;
  }
}

Material default_record() {
  Material m{};
  m.albedo_rgba8 = 0xFFFFFFFFu;
  m.ior = 1.5f;
  m.roughness = 1.0f;
  return m;
}

uint32_t float_bits(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  return u;
}
float bits_float(uint32_t u) {
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}

uint32_t pack_unorm8(const float* v, size_t n) {
  uint32_t packed = 0;
  for (size_t c = 0; c < 4; ++c) {
    const float x = c < n ? v[c] : 1.0f;
    const uint32_t q = uint32_t(std::lround(std::fmin(std::fmax(x, 0.f), 1.f) * 255.f));
    packed |= q << (8 * c);
  }
  return packed;
}

}  // namespace

const char* const_kind_name(int32_t k) { return k >= 0 && k < kConstKinds ? kConst[k].name : "?"; }
uint32_t const_kind_blocks(int32_t k) { return k >= 0 && k < kConstKinds ? kConst[k].blocks : 0; }
const char* shading_input_name(int i) { return i >= 0 && i < kInputs ? kInputNames[i] : "?"; }

MaterialStore::MaterialStore() {
  // The layout of mrt_material, that of the reference application's material type
  // (src/main.cpp:447-464): color_t {value: vec4unorm8}, then ior, roughness, metallic,
  // transmission and albedo_texture (uvec2): 7 blocks. Hidden from name lookups, so a caller may
  // register layouts of the same names.
  uint32_t color = 0, mat = 0, type = 0;
  add_layout("color_t", &color);
  add_parameter(color, "value", "vec4unorm8");
  add_layout("material", &mat);
  add_parameter(mat, "albedo", "color_t");
  add_parameter(mat, "ior", "float");
  add_parameter(mat, "roughness", "float");
  add_parameter(mat, "metallic", "float");
  add_parameter(mat, "transmission", "float");
  add_parameter(mat, "albedo_texture", "uvec2");
  push_material_type(mat, nullptr, &type);
  layouts[color].hidden = layouts[mat].hidden = true;
}

int MaterialStore::add_layout(const std::string& name, uint32_t* out) {
  if (name.empty()) return fail(MRT_ERR_INVALID, "a layout needs a name");
  layouts.push_back(Layout{name, 0, {}, false, false});
  if (out) *out = uint32_t(layouts.size() - 1);
  return MRT_OK;
}

bool MaterialStore::find_type(const std::string& name, TypeRef* out) const {
  for (size_t i = 0; i < layouts.size(); ++i)
    if (!layouts[i].hidden && layouts[i].name == name) {
      *out = TypeRef{-1, uint32_t(i)};
      return true;
    }
  for (int32_t k = 0; k < kConstKinds; ++k)
    if (name == kConst[k].name) {
      *out = TypeRef{k, 0};
      return true;
    }
  return false;
}

int MaterialStore::add_parameter(uint32_t layout, const std::string& name, const std::string& type_name) {
  if (layout >= layouts.size() || layouts[layout].hidden) return fail(MRT_ERR_INVALID, "unknown layout");
  Layout& l = layouts[layout];
  if (l.sealed) return fail(MRT_ERR_INVALID, "layout '" + l.name + "' is in use and cannot grow");
  if (name.empty()) return fail(MRT_ERR_INVALID, "a parameter needs a name");
  TypeRef t;
  if (!find_type(type_name, &t)) return fail(MRT_ERR_INVALID, "unknown parameter type '" + type_name + "'");
  if (t.constant < 0 && t.object == layout) return fail(MRT_ERR_INVALID, "a layout cannot contain itself");
  const uint32_t b = t.constant >= 0 ? kConst[t.constant].blocks : layouts[t.object].blocks;
  if (b == 0) return fail(MRT_ERR_INVALID, "layout '" + type_name + "' has no parameters");
  if (l.blocks + b > kMaxLayoutBlocks) return fail(MRT_ERR_LIMIT, "a layout holds at most 65536 blocks");
  if (t.constant < 0) layouts[t.object].sealed = true;
  // add_parameter (src/types.hpp:250-258): the parameter starts where the previous one ended.
  l.params.push_back(LayoutParam{name, l.blocks, b, t});
  l.blocks += b;
  return MRT_OK;
}

const LayoutParam* MaterialStore::find_param(uint32_t layout, const std::string& name) const {
  for (const LayoutParam& p : layouts[layout].params)
    if (p.name == name) return &p;  // the first of that name, as material_buffer::set finds it
  return nullptr;
}

// The constant kind behind a parameter, descending through layouts of a single parameter (the
// reference's color_t wraps one vec4unorm8); -1 for a structured parameter. *offset accumulates
// the leaf's block offset.
int32_t MaterialStore::leaf_kind(const LayoutParam& p, uint32_t* offset) const {
  const LayoutParam* q = &p;
  *offset += q->offset;
  while (q->type.constant < 0) {
    const Layout& inner = layouts[q->type.object];
    if (inner.params.size() != 1) return -1;
    q = &inner.params[0];
    *offset += q->offset;
  }
  return q->type.constant;
}

int MaterialStore::resolve_path(uint32_t layout, const std::string& path, InputBinding* out) const {
  uint32_t cur = layout, offset = 0;
  size_t start = 0;
  for (;;) {
    const size_t dot = path.find('.', start);
    const std::string part = path.substr(start, dot == std::string::npos ? std::string::npos : dot - start);
    const LayoutParam* p = find_param(cur, part);
    if (!p) return fail(MRT_ERR_INVALID, "layout '" + layouts[cur].name + "' has no parameter '" + part + "'");
    if (dot == std::string::npos) {
      const int32_t k = leaf_kind(*p, &offset);
      if (k < 0) return fail(MRT_ERR_INVALID, "parameter path '" + path + "' names a structured parameter");
      *out = InputBinding{int32_t(offset), k};
      return MRT_OK;
    }
    if (p->type.constant >= 0) return fail(MRT_ERR_INVALID, "parameter '" + part + "' has no members");
    offset += p->offset;
    cur = p->type.object;
    start = dot + 1;
  }
}

int MaterialStore::push_material_type(uint32_t layout, const char* const* paths, uint32_t* out) {
  if (layout >= layouts.size()) return fail(MRT_ERR_INVALID, "unknown layout");
  MaterialType mt;
  mt.layout = layout;
  for (int i = 0; i < kInputs; ++i) {
    InputBinding b;
    if (paths && paths[i]) {
      if (paths[i][0] == '\0') continue;  // explicitly unbound
      int rc = resolve_path(layout, paths[i], &b);
      if (rc != MRT_OK) return rc;
    } else {
      const LayoutParam* p = find_param(layout, kInputNames[i]);
      if (!p) continue;  // the input keeps its default
      uint32_t off = 0;
      const int32_t k = leaf_kind(*p, &off);
      b = InputBinding{int32_t(off), k};
    }
    if (!input_accepts(i, b.kind))
      return fail(MRT_ERR_INVALID, std::string("the ") + kInputNames[i] + " input cannot take a parameter of type " +
                                       (b.kind >= 0 ? kConst[b.kind].name : "structured"));
    mt.in[i] = b;
  }
  layouts[layout].sealed = true;
  types.push_back(mt);
  if (out) *out = uint32_t(types.size() - 1);
  return MRT_OK;
}

int MaterialStore::push_material(uint32_t type, uint32_t* out) {
  if (type >= types.size()) return fail(MRT_ERR_INVALID, "unknown material type");
  const uint32_t b = layouts[types[type].layout].blocks;
  if (blocks.size() + b > kMaxBlocks) return fail(MRT_ERR_LIMIT, "material buffer exceeds 2^30 blocks");
  // material_buffer::push (src/material.hpp:81-91): the new material's blocks start at zero.
  materials.push_back(MaterialSlot{type, uint32_t(blocks.size())});
  blocks.resize(blocks.size() + b, 0u);
  if (out) *out = uint32_t(materials.size() - 1);
  return MRT_OK;
}

int MaterialStore::set_raw(uint32_t material, const std::string& name, const void* data, size_t bytes, bool* found) {
  if (material >= materials.size()) return fail(MRT_ERR_INVALID, "unknown material");
  const MaterialSlot& s = materials[material];
  const LayoutParam* p = find_param(types[s.type].layout, name);
  *found = p != nullptr;
  if (!p) return MRT_OK;
  if (bytes != size_t(p->blocks) * 4) return fail(MRT_ERR_INVALID, "Data size not suitable for this parameter");
  std::memcpy(&blocks[s.offset + p->offset], data, bytes);
  return MRT_OK;
}

int MaterialStore::set_floats(uint32_t material, const std::string& name, const float* v, size_t n,
                              uint32_t num_textures) {
  if (material >= materials.size()) return fail(MRT_ERR_INVALID, "unknown material");
  const MaterialSlot& s = materials[material];
  const MaterialType& mt = types[s.type];
  const LayoutParam* p = find_param(mt.layout, name);
  if (!p) return fail(MRT_ERR_INVALID, "unknown material parameter '" + name + "'");
  uint32_t off = 0;
  const int32_t k = leaf_kind(*p, &off);
  if (k < 0) return fail(MRT_ERR_INVALID, "parameter '" + name + "' is structured: set it from its bytes");
  uint32_t* dst = &blocks[s.offset + off];
  if (k == kVec4Unorm8) {
    if (n != 3 && n != 4) return fail(MRT_ERR_INVALID, "parameter '" + name + "' (vec4unorm8) takes 3 or 4 values");
    dst[0] = pack_unorm8(v, n);
    return MRT_OK;
  }
  const uint32_t b = kConst[k].blocks;
  const bool texture = mt.in[kInTexture].offset == int32_t(off);
  if (n != b && !(texture && n == 1))  // a texture input also takes its id alone
    return fail(MRT_ERR_INVALID, "parameter '" + name + "' takes " + std::to_string(b) + " values");
  if (!is_integer_kind(k)) {
    for (size_t i = 0; i < n; ++i) dst[i] = float_bits(v[i]);
    return MRT_OK;
  }
  for (size_t i = 0; i < n; ++i) {
    const float x = v[i];
    if (!(x == std::floor(x)) || std::fabs(x) > 2147483647.f || (is_unsigned_kind(k) && x < 0.f))
      return fail(MRT_ERR_INVALID, "parameter '" + name + "' takes integral values");
    if (texture && i == 0 && x > float(num_textures))
      return fail(MRT_ERR_INVALID, "parameter '" + name + "' is not a texture id");
  }
  for (size_t i = 0; i < b; ++i) dst[i] = i < n ? uint32_t(int64_t(v[i])) : 0u;
  return MRT_OK;
}

int MaterialStore::set_record(uint32_t material, const Material& m) {
  if (material >= materials.size()) return fail(MRT_ERR_INVALID, "unknown material");
  const MaterialSlot& s = materials[material];
  const MaterialType& mt = types[s.type];
  const Material def = default_record();
  const uint32_t vals[kInputs] = {m.albedo_rgba8, float_bits(m.ior), float_bits(m.roughness), float_bits(m.metallic),
                                  float_bits(m.transmission), m.texture};
  const uint32_t defs[kInputs] = {def.albedo_rgba8, float_bits(def.ior), float_bits(def.roughness),
                                  float_bits(def.metallic), float_bits(def.transmission), def.texture};
  for (int i = 0; i < kInputs; ++i)
    if (mt.in[i].offset < 0 && vals[i] != defs[i])
      return fail(MRT_ERR_INVALID, std::string("material type has no parameter for the ") + kInputNames[i] + " input");
  for (int i = 0; i < kInputs; ++i) {
    if (mt.in[i].offset < 0) continue;
    uint32_t* dst = &blocks[s.offset + uint32_t(mt.in[i].offset)];
    dst[0] = vals[i];
    if (mt.in[i].kind == kUVec2) dst[1] = 0u;
  }
  return MRT_OK;
}

Material MaterialStore::resolve(uint32_t material) const {
  Material m = default_record();
  const MaterialSlot& s = materials[material];
  const MaterialType& mt = types[s.type];
  const auto at = [&](int i) { return blocks[s.offset + uint32_t(mt.in[i].offset)]; };
  if (mt.in[kInAlbedo].offset >= 0) m.albedo_rgba8 = at(kInAlbedo);
  if (mt.in[kInIor].offset >= 0) m.ior = bits_float(at(kInIor));
  if (mt.in[kInRoughness].offset >= 0) m.roughness = bits_float(at(kInRoughness));
  if (mt.in[kInMetallic].offset >= 0) m.metallic = bits_float(at(kInMetallic));
  if (mt.in[kInTransmission].offset >= 0) m.transmission = bits_float(at(kInTransmission));
  if (mt.in[kInTexture].offset >= 0) {
    // A uvec2 handle names one of this scene's textures in its low word; any other handle (a
    // nonzero high word) is no texture here. Ids past the last texture are ignored by the shader.
    const uint32_t lo = at(kInTexture);
    const bool wide = mt.in[kInTexture].kind == kUVec2;
    m.texture = wide && blocks[s.offset + uint32_t(mt.in[kInTexture].offset) + 1] != 0u ? 0u : lo;
  }
  return m;
}

}  // namespace mrt

using namespace mrt;

namespace {
Scene* scene_of(mrt_scene h) { return reinterpret_cast<Scene*>(h); }
}  // namespace

extern "C" {

int mrt_scene_push_material(mrt_scene sh, const mrt_material* mat, uint32_t* out) {
  MRT_GUARD_BEGIN
  Scene* s = scene_of(sh);
  if (!s || !mat) return fail(MRT_ERR_INVALID, "null argument");
  static_assert(sizeof(mrt_material) == sizeof(Material), "material layouts differ");
  Material m;
  std::memcpy(&m, mat, sizeof(m));
  uint32_t id = 0;
  int rc = s->mats.push_material(MaterialStore::kRecordType, &id);
  if (rc == MRT_OK) rc = s->mats.set_record(id, m);
  if (rc != MRT_OK) return rc;
  s->shading_dirty = true;
  if (out) *out = id;
  return MRT_OK;
  MRT_GUARD_END
}

int mrt_scene_set_material(mrt_scene sh, uint32_t id, const mrt_material* mat) {
  MRT_GUARD_BEGIN
  Scene* s = scene_of(sh);
  if (!s || !mat) return fail(MRT_ERR_INVALID, "null argument");
  Material m;
  std::memcpy(&m, mat, sizeof(m));
  const int rc = s->mats.set_record(id, m);
  if (rc == MRT_OK) s->shading_dirty = true;
  return rc;
  MRT_GUARD_END
}

int mrt_scene_get_material(mrt_scene sh, uint32_t id, mrt_material* out) {
  MRT_GUARD_BEGIN
  Scene* s = scene_of(sh);
  if (!s || !out) return fail(MRT_ERR_INVALID, "null argument");
  if (id >= s->mats.materials.size()) return fail(MRT_ERR_INVALID, "unknown material");
  const Material m = s->mats.resolve(id);
  std::memcpy(out, &m, sizeof(m));
  return MRT_OK;
  MRT_GUARD_END
}

int mrt_scene_set_material_param(mrt_scene sh, uint32_t id, const char* name, const float* value, size_t n) {
  MRT_GUARD_BEGIN
  Scene* s = scene_of(sh);
  if (!s || !name || (n && !value)) return fail(MRT_ERR_INVALID, "null argument");
  const int rc = s->mats.set_floats(id, name, value, n, uint32_t(s->textures.size()));
  if (rc == MRT_OK) s->shading_dirty = true;
  return rc;
  MRT_GUARD_END
}

int mrt_scene_register_layout(mrt_scene sh, const char* name, uint32_t* out_layout) {
  MRT_GUARD_BEGIN
  Scene* s = scene_of(sh);
  if (!s || !name) return fail(MRT_ERR_INVALID, "null argument");
  return s->mats.add_layout(name, out_layout);
  MRT_GUARD_END
}

int mrt_scene_layout_add_parameter(mrt_scene sh, uint32_t layout, const char* name, const char* type_name) {
  MRT_GUARD_BEGIN
  Scene* s = scene_of(sh);
  if (!s || !name || !type_name) return fail(MRT_ERR_INVALID, "null argument");
  return s->mats.add_parameter(layout, name, type_name);
  MRT_GUARD_END
}

int mrt_scene_find_type(mrt_scene sh, const char* name, int32_t* constant, uint32_t* layout) {
  MRT_GUARD_BEGIN
  Scene* s = scene_of(sh);
  if (!s || !name) return fail(MRT_ERR_INVALID, "null argument");
  TypeRef t;
  if (!s->mats.find_type(name, &t)) return fail(MRT_ERR_INVALID, std::string("unknown type '") + name + "'");
  if (constant) *constant = t.constant;
  if (layout) *layout = t.constant >= 0 ? 0xFFFFFFFFu : t.object;
  return MRT_OK;
  MRT_GUARD_END
}

int mrt_scene_layout_info(mrt_scene sh, uint32_t layout, uint32_t* blocks, uint32_t* n_params) {
  MRT_GUARD_BEGIN
  Scene* s = scene_of(sh);
  if (!s) return fail(MRT_ERR_INVALID, "null argument");
  if (layout >= s->mats.layouts.size()) return fail(MRT_ERR_INVALID, "unknown layout");
  const Layout& l = s->mats.layouts[layout];
  if (blocks) *blocks = l.blocks;
  if (n_params) *n_params = uint32_t(l.params.size());
  return MRT_OK;
  MRT_GUARD_END
}

int mrt_scene_layout_parameter(mrt_scene sh, uint32_t layout, uint32_t index, char* name, size_t name_cap,
                               uint32_t* offset, uint32_t* blocks, int32_t* constant, uint32_t* inner_layout) {
  MRT_GUARD_BEGIN
  Scene* s = scene_of(sh);
  if (!s) return fail(MRT_ERR_INVALID, "null argument");
  if (layout >= s->mats.layouts.size()) return fail(MRT_ERR_INVALID, "unknown layout");
  const Layout& l = s->mats.layouts[layout];
  if (index >= l.params.size()) return fail(MRT_ERR_INVALID, "parameter index out of range");
  const LayoutParam& p = l.params[index];
  if (name && name_cap) {
    if (p.name.size() + 1 > name_cap) return fail(MRT_ERR_LIMIT, "name buffer too small");
    std::memcpy(name, p.name.c_str(), p.name.size() + 1);
  }
  if (offset) *offset = p.offset;
  if (blocks) *blocks = p.blocks;
  if (constant) *constant = p.type.constant;
  if (inner_layout) *inner_layout = p.type.constant >= 0 ? 0xFFFFFFFFu : p.type.object;
  return MRT_OK;
  MRT_GUARD_END
}

int mrt_scene_push_material_type(mrt_scene sh, uint32_t layout, const char* const* input_paths,
                                 uint32_t* out_type) {
  MRT_GUARD_BEGIN
  Scene* s = scene_of(sh);
  if (!s) return fail(MRT_ERR_INVALID, "null argument");
  if (layout < s->mats.layouts.size() && s->mats.layouts[layout].hidden)
    return fail(MRT_ERR_INVALID, "unknown layout");
  return s->mats.push_material_type(layout, input_paths, out_type);
  MRT_GUARD_END
}

int mrt_scene_push_material_of_type(mrt_scene sh, uint32_t type, uint32_t* out) {
  MRT_GUARD_BEGIN
  Scene* s = scene_of(sh);
  if (!s) return fail(MRT_ERR_INVALID, "null argument");
  const int rc = s->mats.push_material(type, out);
  if (rc == MRT_OK) s->shading_dirty = true;
  return rc;
  MRT_GUARD_END
}

int mrt_scene_set_parameter(mrt_scene sh, uint32_t material, const char* name, const void* data, size_t bytes,
                            int* found) {
  MRT_GUARD_BEGIN
  Scene* s = scene_of(sh);
  if (!s || !name || (bytes && !data)) return fail(MRT_ERR_INVALID, "null argument");
  bool f = false;
  const int rc = s->mats.set_raw(material, name, data, bytes, &f);
  if (found) *found = f ? 1 : 0;
  if (rc == MRT_OK && f) s->shading_dirty = true;
  return rc;
  MRT_GUARD_END
}

int mrt_scene_material_slot(mrt_scene sh, uint32_t material, uint32_t* type, uint32_t* offset) {
  MRT_GUARD_BEGIN
  Scene* s = scene_of(sh);
  if (!s) return fail(MRT_ERR_INVALID, "null argument");
  if (material >= s->mats.materials.size()) return fail(MRT_ERR_INVALID, "unknown material");
  if (type) *type = s->mats.materials[material].type;
  if (offset) *offset = s->mats.materials[material].offset;
  return MRT_OK;
  MRT_GUARD_END
}

int mrt_scene_material_blocks(mrt_scene sh, uint32_t* out, size_t n, size_t* total) {
  MRT_GUARD_BEGIN
  Scene* s = scene_of(sh);
  if (!s || (n && !out)) return fail(MRT_ERR_INVALID, "null argument");
  const std::vector<uint32_t>& b = s->mats.blocks;
  if (total) *total = b.size();
  if (n) std::memcpy(out, b.data(), std::min(n, b.size()) * 4);
  return MRT_OK;
  MRT_GUARD_END
}

}  // extern "C"
