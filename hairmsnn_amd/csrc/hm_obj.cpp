// Wavefront OBJ head meshes (loadOBJ, model.cpp:233-358, which reads them
// with tinyobjloader, triangulating): positions, vertex normals and texture
// coordinates, polygons fanned into triangles, and per-face materials from
// the referenced .mtl file (Kd as the diffuse colour, 0 when absent). Every
// triangle keeps its own copy of its three vertices, as the reference's
// per-material meshes do. Textured materials (map_Kd, map_Ns) are refused
// with a message: no image decoder is built.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/hmsnn.h"
#include "hm_api_util.h"
#include "hm_scene.h"

namespace hm {

namespace {

struct ObjMaterial {
  V3 kd{0.f, 0.f, 0.f};
};

std::string dir_part(const std::string& path) {
  const size_t k = path.find_last_of("/\\");
  return k == std::string::npos ? std::string() : path.substr(0, k + 1);
}

int parse_mtl(const std::string& path, std::map<std::string, ObjMaterial>& mats, std::string& err) {
  std::ifstream f(path);
  if (!f) {
    err = "Could not read OBJ material library " + path;
    return HMSNN_EIO;
  }
  std::string line, cur;
  while (std::getline(f, line)) {
    std::istringstream ls(line);
    std::string key;
    if (!(ls >> key) || key[0] == '#') continue;
    if (key == "newmtl") {
      ls >> cur;
      mats[cur] = ObjMaterial{};
    } else if (key == "Kd" && !cur.empty()) {
      float r = 0, g = 0, b = 0;
      ls >> r >> g >> b;
      mats[cur].kd = v3(r, g, b);
    } else if (key == "map_Kd" || key == "map_Ns") {
      err = path + ": textured materials (" + key + ") are not supported";
      return HMSNN_EUNSUPPORTED;
    }
  }
  return HMSNN_OK;
}

// One "v[/vt][/vn]" reference; indices are 1-based, negative counts back.
bool parse_ref(const std::string& tok, int nv, int nvt, int nvn, int& vi, int& ti, int& ni) {
  int idx[3] = {0, 0, 0};
  size_t pos = 0;
  for (int k = 0; k < 3 && pos <= tok.size(); ++k) {
    const size_t slash = tok.find('/', pos);
    const std::string part = tok.substr(pos, slash == std::string::npos ? std::string::npos : slash - pos);
    if (!part.empty()) idx[k] = std::atoi(part.c_str());
    if (slash == std::string::npos) break;
    pos = slash + 1;
  }
  auto fix = [](int i, int n) { return i > 0 ? i - 1 : (i < 0 ? n + i : -1); };
  vi = fix(idx[0], nv);
  ti = fix(idx[1], nvt);
  ni = fix(idx[2], nvn);
  return vi >= 0 && vi < nv && ti < nvt && ni < nvn;
}

}  // namespace

int load_obj(const std::string& path, TriMesh& out, std::string& err) {
  std::ifstream f(path);
  if (!f) {
    err = "Could not read OBJ model from " + path;
    return HMSNN_EIO;
  }
  out = TriMesh{};
  std::vector<V3> pos, nrm;
  int nvt = 0;
  std::map<std::string, ObjMaterial> mats;
  V3 kd = v3(0.f);
  std::string line;
  int line_no = 0;
  while (std::getline(f, line)) {
    ++line_no;
    std::istringstream ls(line);
    std::string key;
    if (!(ls >> key) || key[0] == '#') continue;
    if (key == "v") {
      float x = 0, y = 0, z = 0;
      ls >> x >> y >> z;
      pos.push_back(v3(x, y, z));
    } else if (key == "vn") {
      float x = 0, y = 0, z = 0;
      ls >> x >> y >> z;
      nrm.push_back(v3(x, y, z));
    } else if (key == "vt") {
      ++nvt;
    } else if (key == "mtllib") {
      std::string lib;
      ls >> lib;
      const int rc = parse_mtl(dir_part(path) + lib, mats, err);
      if (rc) return rc;
    } else if (key == "usemtl") {
      std::string name;
      ls >> name;
      auto it = mats.find(name);
      kd = it == mats.end() ? v3(0.f) : it->second.kd;
    } else if (key == "f") {
      std::vector<int> vi, ni;
      std::string tok;
      while (ls >> tok) {
        int a, b, c;
        if (!parse_ref(tok, (int)pos.size(), nvt, (int)nrm.size(), a, b, c)) {
          err = path + ":" + std::to_string(line_no) + ": invalid face index '" + tok + "'";
          return HMSNN_EIO;
        }
        vi.push_back(a);
        ni.push_back(c);
      }
      if (vi.size() < 3) {
        err = path + ":" + std::to_string(line_no) + ": face with fewer than 3 vertices";
        return HMSNN_EIO;
      }
      for (size_t k = 1; k + 1 < vi.size(); ++k) {  // fan triangulation
        const int t[3] = {0, (int)k, (int)k + 1};
        const V3 p0 = pos[vi[0]], p1 = pos[vi[k]], p2 = pos[vi[k + 1]];
        const V3 fn = normalize(cross(p1 - p0, p2 - p0));
        for (int j = 0; j < 3; ++j) {
          const V3 p = pos[vi[t[j]]];
          out.v.insert(out.v.end(), {p.x, p.y, p.z});
          const V3 n = ni[t[j]] >= 0 ? nrm[ni[t[j]]] : fn;  // face normal when the file has none
          out.n.insert(out.n.end(), {n.x, n.y, n.z});
          out.bmin = vmin(out.bmin, p);
          out.bmax = vmax(out.bmax, p);
        }
        out.kd.insert(out.kd.end(), {kd.x, kd.y, kd.z});
      }
    }
  }
  if (out.v.empty()) {
    err = "OBJ model " + path + " has no faces";
    return HMSNN_EIO;
  }
  return HMSNN_OK;
}

}  // namespace hm

extern "C" int hmsnn_obj_load(const char* path, int* num_triangles, float* vertices, float* normals, float* diffuse,
                              int capacity) {
  if (!path || !num_triangles) return hm::fail(HMSNN_EINVAL, "hmsnn_obj_load: bad arguments");
  hm::TriMesh m;
  std::string err;
  const int rc = hm::load_obj(path, m, err);
  if (rc) return hm::fail(rc, err);
  *num_triangles = m.num();
  if (vertices || normals || diffuse) {
    if (capacity < m.num()) return hm::fail(HMSNN_EINVAL, "hmsnn_obj_load: capacity below the triangle count");
    if (vertices) std::memcpy(vertices, m.v.data(), sizeof(float) * m.v.size());
    if (normals) std::memcpy(normals, m.n.data(), sizeof(float) * m.n.size());
    if (diffuse) std::memcpy(diffuse, m.kd.data(), sizeof(float) * m.kd.size());
  }
  return HMSNN_OK;
}
