// Wavefront OBJ/MTL ingest: myrt::obj::load_mtllib, load_obj and triangulate (src/obj.cpp:9-361)
// and the app's load_object_file (src/main.cpp:902-949), which pushes every triangulated group
// into a scene as a geometry with a material mapped from its MTL entry.
//
// Parsing follows the reference statement by statement; where the reference reads out of bounds
// or dereferences a null material (a face before any object, `usemtl` of an unknown name, a face
// vertex without slashes, an index beyond its list), this parser defines the result instead:
// the object is created, the default material is used, a missing index field is 0, and an index
// out of range is an error.
#include <array>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <memory>
#include <sstream>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/myrt.h"
#include "scene_internal.hpp"
#include "abi_guard.hpp"

struct mrt_obj_s;

namespace mrt {
namespace {

struct ObjMaterial {  // myrt::obj::material_t with make_default_material's values (obj.cpp:84-95)
  std::string name = "Default";
  float specularity = 100.f;  // Ns
  float diffuse[3] = {1.f, 0.f, 1.f};
  float ior = 1.45f;  // Ni
  float dissolve = 1.f;  // d
  std::string map_diffuse;
};
using MaterialPtr = std::shared_ptr<ObjMaterial>;
using MaterialLib = std::unordered_map<std::string, MaterialPtr>;

struct ObjGroup {
  std::string name;
  std::vector<std::vector<std::array<long, 3>>> faces;  // per face vertex: position, uv, normal index
  MaterialPtr material;
};

struct ObjObject {
  std::string name;
  std::vector<ObjGroup> groups;
  std::vector<std::array<float, 3>> positions, normals;
  std::vector<std::array<float, 2>> texcoords;
};

std::string parent_dir(const std::string& path) {
  const size_t k = path.find_last_of('/');
  return k == std::string::npos ? std::string() : path.substr(0, k + 1);
}

// load_mtllib (obj.cpp:9-80): statements before the first newmtl are ignored here; Ka, Ks, Ke
// and illum are not used by the material mapping and are not kept.
MaterialLib load_mtllib(const std::string& path) {
  MaterialLib lib;
  std::ifstream in(path);
  MaterialPtr cur;
  for (std::string line; std::getline(in, line);) {
    std::istringstream ls(line);
    std::string id;
    ls >> id;
    if (id == "newmtl") {
      std::string name;
      ls >> name;
      cur = lib[name] = std::make_shared<ObjMaterial>();
      cur->name = name;
    } else if (!cur) {
      continue;
    } else if (id == "Ns") {
      ls >> cur->specularity;
    } else if (id == "Ni") {
      ls >> cur->ior;
    } else if (id == "d") {
      ls >> cur->dissolve;
    } else if (id == "Kd") {
      ls >> cur->diffuse[0] >> cur->diffuse[1] >> cur->diffuse[2];
    } else if (id == "map_Kd") {
      ls >> cur->map_diffuse;
    }
  }
  return lib;
}

// One face vertex "p", "p/t", "p//n" or "p/t/n" (obj.cpp:217-243): each field by strtol, an
// absent field is 0. A negative index counts back from the end of the object's list, a positive
// one is 1-based over the whole file (offset = 1 + the sizes of the earlier objects' lists).
std::array<long, 3> parse_vertex(const std::string& tok, const ObjObject& o, const size_t offset[3]) {
  long f[3] = {0, 0, 0};
  const char* p = tok.c_str();
  for (int k = 0; k < 3; ++k) {
    char* end = nullptr;
    f[k] = std::strtol(p, &end, 10);
    if (*end != '/') break;
    p = end + 1;
  }
  const size_t sizes[3] = {o.positions.size(), o.texcoords.size(), o.normals.size()};
  std::array<long, 3> v{};
  for (int k = 0; k < 3; ++k)
    v[k] = f[k] < 0 ? long(sizes[k]) + f[k] : (f[k] != 0 ? f[k] - long(offset[k]) : 0);
  return v;
}

// load_obj (obj.cpp:96-271).
bool load_obj(const std::string& path, std::vector<ObjObject>& result, std::string& err) {
  std::ifstream in(path, std::ios::binary);
  if (!in) return err = "cannot open '" + path + "'", false;
  std::stringstream text;
  text << in.rdbuf();
  size_t offset[3] = {1, 1, 1};  // positions, texcoords, normals
  std::vector<std::array<float, 3>> pend_pos, pend_nrm;
  std::vector<std::array<float, 2>> pend_tex;
  MaterialLib lib;
  const MaterialPtr fallback = std::make_shared<ObjMaterial>();
  MaterialPtr material = fallback;
  auto current = [&]() -> ObjObject& {
    if (result.empty()) result.emplace_back();
    return result.back();
  };
  for (std::string line; std::getline(text, line);) {
    std::istringstream ls(line);
    std::string id;
    ls >> id;
    if (id == "mtllib") {
      // Every named library is read; a name defined by an earlier library keeps that definition.
      lib.clear();
      for (std::string file; ls >> file;) {
        const MaterialLib more = load_mtllib(parent_dir(path) + file);
        lib.insert(more.begin(), more.end());
      }
    } else if (id == "usemtl") {
      std::string name;
      ls >> name;
      const auto it = lib.find(name);
      material = it != lib.end() ? it->second : fallback;
      ObjObject& o = current();
      if (o.groups.empty() || !o.groups.back().faces.empty()) o.groups.emplace_back().name = name;
      o.groups.back().material = material;
    } else if (id == "o") {
      if (!result.empty()) {
        offset[0] += result.back().positions.size();
        offset[1] += result.back().texcoords.size();
        offset[2] += result.back().normals.size();
      }
      result.emplace_back();
      ls >> result.back().name;
    } else if (id == "g") {
      std::string name;
      ls >> name;
      if (result.empty()) result.emplace_back().name = name;
      ObjGroup& g = result.back().groups.emplace_back();
      g.name = name;
      g.material = material;
    } else if (id == "v" || id == "vn") {
      std::array<float, 3> v{0.f, 0.f, 0.f};
      ls >> v[0] >> v[1] >> v[2];
      (id == "v" ? pend_pos : pend_nrm).push_back(v);
    } else if (id == "vt") {
      std::array<float, 2> v{0.f, 0.f};
      ls >> v[0] >> v[1];
      pend_tex.push_back(v);
    } else if (id == "f") {
      ObjObject& o = current();
      if (o.groups.empty()) {
        ObjGroup& g = o.groups.emplace_back();
        g.name = "Default";
        g.material = material;
      }
      // Vertex data read since the last face belongs to the current object.
      o.positions.insert(o.positions.end(), pend_pos.begin(), pend_pos.end());
      o.normals.insert(o.normals.end(), pend_nrm.begin(), pend_nrm.end());
      o.texcoords.insert(o.texcoords.end(), pend_tex.begin(), pend_tex.end());
      pend_pos.clear();
      pend_nrm.clear();
      pend_tex.clear();
      auto& face = o.groups.back().faces.emplace_back();
      for (std::string tok; ls >> tok;) face.push_back(parse_vertex(tok, o, offset));
    }
  }
  for (ObjObject& o : result) {  // obj.cpp:262-268
    if (o.texcoords.empty()) o.texcoords.push_back({0.f, 0.f});
    if (o.normals.empty()) o.normals.push_back({0.f, 1.f, 0.f});
  }
  return true;
}

// A vertex key compares as floats compare: 0 and -0 are one key, a NaN matches nothing.
struct VertexKey {
  float v[8];
  bool operator==(const VertexKey& o) const {
    for (int k = 0; k < 8; ++k)
      if (!(v[k] == o.v[k])) return false;
    return true;
  }
};
struct VertexKeyHash {
  size_t operator()(const VertexKey& k) const {
    uint64_t h = 0x9e3779b97f4a7c15ull;
    for (float f : k.v) {
      uint32_t b;
      f = f == 0.f ? 0.f : f;
      std::memcpy(&b, &f, 4);
      h = (h ^ b) * 0x100000001b3ull;
      h ^= h >> 29;
    }
    return size_t(h);
  }
};

}  // namespace

struct ObjMesh {  // myrt::obj::triangulated_object_t
  std::string name;
  std::vector<uint32_t> indices;
  std::vector<float> positions, normals, texcoords;
  MaterialPtr material;
};

namespace {

// triangulate (obj.cpp:297-361): one mesh per group, named "object/group". Vertices are shared
// per distinct (position, normal, uv); a face of k >= 3 vertices becomes the fan
// (v0, v_{i-1}, v_i), i = 2..k-1; faces of fewer vertices are dropped.
bool triangulate(const ObjObject& o, std::vector<ObjMesh>& out, std::string& err) {
  for (const ObjGroup& g : o.groups) {
    ObjMesh& m = out.emplace_back();
    m.name = o.name + "/" + g.name;
    m.material = g.material;
    std::unordered_map<VertexKey, uint32_t, VertexKeyHash> seen;
    auto vertex = [&](const std::array<long, 3>& v, uint32_t& index) {
      if (v[0] < 0 || size_t(v[0]) >= o.positions.size() || v[1] < 0 || size_t(v[1]) >= o.texcoords.size() ||
          v[2] < 0 || size_t(v[2]) >= o.normals.size())
        return false;
      const auto& p = o.positions[v[0]];
      const auto& t = o.texcoords[v[1]];
      const auto& n = o.normals[v[2]];
      const VertexKey key{{p[0], p[1], p[2], n[0], n[1], n[2], t[0], t[1]}};
      const auto [it, added] = seen.emplace(key, uint32_t(m.positions.size() / 3));
      if (added) {
        m.positions.insert(m.positions.end(), p.begin(), p.end());
        m.normals.insert(m.normals.end(), n.begin(), n.end());
        m.texcoords.insert(m.texcoords.end(), t.begin(), t.end());
      }
      index = it->second;
      return true;
    };
    for (const auto& face : g.faces) {
      if (face.size() <= 2) continue;
      uint32_t first, last;
      bool ok = vertex(face[0], first) && vertex(face[1], last);
      for (size_t i = 2; ok && i < face.size(); ++i) {
        uint32_t next;
        ok = vertex(face[i], next);
        if (ok) m.indices.insert(m.indices.end(), {first, last, next});
        last = next;
      }
      if (!ok) return err = "face index out of range in object '" + o.name + "', group '" + g.name + "'", false;
    }
  }
  return true;
}

void copy_name(char* dst, size_t cap, const std::string& s) {
  const size_t n = std::min(cap - 1, s.size());
  std::memcpy(dst, s.data(), n);
  dst[n] = 0;
}

}  // namespace
}  // namespace mrt

struct mrt_obj_s {
  std::vector<mrt::ObjMesh> meshes;
  std::string dir;  // the OBJ's folder, for map_Kd paths
};

using namespace mrt;


extern "C" {

int mrt_obj_load(const char* path, mrt_obj* out) {
  MRT_GUARD_BEGIN
  if (!path || !out) return fail(MRT_ERR_INVALID, "null argument");
  *out = nullptr;
  std::vector<ObjObject> objects;
  std::string err;
  if (!load_obj(path, objects, err)) return fail(MRT_ERR_INVALID, err);
  auto o = std::make_unique<mrt_obj_s>();
  o->dir = parent_dir(path);
  for (const ObjObject& obj : objects)
    if (!triangulate(obj, o->meshes, err)) return fail(MRT_ERR_INVALID, err + " of '" + path + "'");
  for (const ObjMesh& m : o->meshes)
    if (m.positions.size() / 3 > 0xFFFFFFFFull) return fail(MRT_ERR_LIMIT, "group has more than 2^32 vertices");
  *out = o.release();
  return MRT_OK;
  MRT_GUARD_END
}

int mrt_obj_destroy(mrt_obj o) {
  MRT_GUARD_BEGIN
  delete o;
  return MRT_OK;
  MRT_GUARD_END
}

int mrt_obj_num_groups(mrt_obj o, size_t* n) {
  MRT_GUARD_BEGIN
  if (!o || !n) return fail(MRT_ERR_INVALID, "null argument");
  *n = o->meshes.size();
  return MRT_OK;
  MRT_GUARD_END
}

int mrt_obj_group(mrt_obj o, size_t i, mrt_obj_group_info* info) {
  MRT_GUARD_BEGIN
  if (!o || !info) return fail(MRT_ERR_INVALID, "null argument");
  if (i >= o->meshes.size()) return fail(MRT_ERR_INVALID, "group index out of range");
  const ObjMesh& m = o->meshes[i];
  const ObjMaterial& mat = *m.material;
  std::memset(info, 0, sizeof(*info));
  info->n_idx = m.indices.size();
  info->n_vtx = m.positions.size() / 3;
  for (int c = 0; c < 3; ++c) info->diffuse[c] = mat.diffuse[c];
  info->specularity = mat.specularity;
  info->ior = mat.ior;
  info->dissolve = mat.dissolve;
  copy_name(info->name, sizeof(info->name), m.name);
  copy_name(info->material_name, sizeof(info->material_name), mat.name);
  if (!mat.map_diffuse.empty()) {
    const std::string tex = mat.map_diffuse[0] == '/' ? mat.map_diffuse : o->dir + mat.map_diffuse;
    if (tex.size() >= sizeof(info->albedo_texture)) return fail(MRT_ERR_LIMIT, "map_Kd path too long");
    copy_name(info->albedo_texture, sizeof(info->albedo_texture), tex);
  }
  return MRT_OK;
  MRT_GUARD_END
}

int mrt_obj_group_copy(mrt_obj o, size_t i, uint32_t* idx, float* pos3, float* nrm3, float* uv2) {
  MRT_GUARD_BEGIN
  if (!o) return fail(MRT_ERR_INVALID, "null argument");
  if (i >= o->meshes.size()) return fail(MRT_ERR_INVALID, "group index out of range");
  const ObjMesh& m = o->meshes[i];
  if (idx) std::memcpy(idx, m.indices.data(), m.indices.size() * 4);
  if (pos3) std::memcpy(pos3, m.positions.data(), m.positions.size() * 4);
  if (nrm3) std::memcpy(nrm3, m.normals.data(), m.normals.size() * 4);
  if (uv2) std::memcpy(uv2, m.texcoords.data(), m.texcoords.size() * 4);
  return MRT_OK;
  MRT_GUARD_END
}

int mrt_scene_push_obj(mrt_scene sh, mrt_obj o, uint32_t* geometries, uint32_t* materials, size_t n) {
  MRT_GUARD_BEGIN
  if (!sh || !o || (o->meshes.size() && (!geometries || !materials))) return fail(MRT_ERR_INVALID, "null argument");
  if (n != o->meshes.size()) return fail(MRT_ERR_INVALID, "n must equal the number of groups");
  for (size_t i = 0; i < n; ++i) {
    const ObjMesh& mesh = o->meshes[i];
    geometries[i] = materials[i] = 0xFFFFFFFFu;
    if (mesh.indices.empty()) continue;  // main.cpp:912-913
    int rc = mrt_scene_push_geometry(sh, mesh.indices.data(), mesh.indices.size(), mesh.positions.data(),
                                     mesh.normals.data(), mesh.texcoords.data(), mesh.positions.size() / 3,
                                     &geometries[i]);
    if (rc != MRT_OK) return rc;
    // main.cpp:919-927: albedo = vec4ui8(255 * (Kd, 1)), converted per component by truncation;
    // roughness = max(0.054, (1 / Ns)^(1 / pi)); transmission = 1 - d; metallic stays 0.
    const ObjMaterial& om = *mesh.material;
    auto unorm = [](float v) { return uint32_t(uint8_t(std::fmin(std::fmax(255.f * v, 0.f), 255.f))); };
    mrt_material m{};
    m.albedo_rgba8 = unorm(om.diffuse[0]) | unorm(om.diffuse[1]) << 8 | unorm(om.diffuse[2]) << 16 | 255u << 24;
    m.ior = om.ior;
    m.roughness = std::max(0.054f, std::pow(1.f / om.specularity, 1.f / 3.1415926535897f));
    m.transmission = 1.0f - om.dissolve;
    if ((rc = mrt_scene_push_material(sh, &m, &materials[i])) != MRT_OK) return rc;
  }
  return MRT_OK;
  MRT_GUARD_END
}

}  // extern "C"
