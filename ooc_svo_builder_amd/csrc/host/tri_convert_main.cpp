// tri_convert / tri_convert_binary: model file -> .tri header + .tridata records, the
// input of svo_builder(_binary). Replaces tri_convert.cpp:85-150 without trimesh2:
// PLY (ascii, binary little/big endian) and Wavefront OBJ are read here.
//
// What the reference does, kept here (tri_convert.cpp:85-150, tri_convert_util.h:20-52):
//   * the mesh bbox over all vertices, padded to a cube around its centre
//     (createMeshBBCube: each short side grows by (maxlength - length) / 2 on both ends);
//   * every vertex translated by -bbox.min, so the grid origin is 0;
//   * the .tri header bbox is the padded cube *before* the translation (tri_convert.cpp:139);
//   * payload records: per-vertex colours when the file has them (else zero), and the face
//     normal = normalize(mean of the 3 vertex normals) (getShadingFaceNormal), or with -r
//     normalize(cross(v0 - v1, v1 - v2)) of the translated vertices (computeFaceNormal).
// Reader behaviour of trimesh2 restated here (trimesh2 is third-party and not part of the
// reference): polygons are split into triangle fans; PLY colours stored as uchar become
// value / 255; a payload mesh without vertex normals gets them from its faces with Max's
// weights (each face normal cross(p0 - p1, p1 - p2) added to a corner scaled by
// 1 / (|edge a|^2 |edge b|^2) of the corner's two edges), then normalized.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#ifndef TRI_PAYLOAD
#define TRI_PAYLOAD 1
#endif

using namespace std;

namespace {

struct V3 {
  float x = 0, y = 0, z = 0;
};
[[maybe_unused]] V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
[[maybe_unused]] V3 cross(V3 a, V3 b) {
  return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
// glm::normalize: v * inversesqrt(dot(v, v)), dot = (x*x + y*y) + z*z
[[maybe_unused]] V3 normalize(V3 v) {
  const float inv = 1.0f / sqrtf((v.x * v.x + v.y * v.y) + v.z * v.z);
  return {v.x * inv, v.y * inv, v.z * inv};
}

struct Mesh {
  vector<V3> vertices, normals, colors;
  vector<uint32_t> faces;  // 3 per triangle
};

string version = "1.6.4";
string filename;
bool recompute_normals = false;

void printInfo() {
  cout << "-------------------------------------------------------------" << endl;
#if TRI_PAYLOAD
  cout << "Tri Converter " << version << " - GEOMETRY+NORMALS" << endl;
#else
  cout << "Tri Converter " << version << " - BINARY VOXELIZATION" << endl;
#endif
#ifdef __linux__
  cout << "Linux ";
#endif
  cout << endl;
  cout << "MI355X build (PLY and OBJ input)" << endl;
  cout << "-------------------------------------------------------------" << endl << endl;
}

void printHelp() {
  cout << "Example: tri_convert -f /home/jeroen/bunny.ply" << endl;
  cout << "" << endl;
  cout << "All available program options:" << endl;
  cout << "" << endl;
  cout << "-f <filename>         Path to a model input file (.ply or .obj)." << endl;
  cout << "-r                    Recompute face normals." << endl;
  cout << "-h                    Print help and exit." << endl;
}

void printInvalid() {
  cout << "Not enough or invalid arguments, please try again.\n" << endl;
  printHelp();
}

// ------------------------------------------------------------------------------ PLY
enum PType { P_CHAR, P_UCHAR, P_SHORT, P_USHORT, P_INT, P_UINT, P_FLOAT, P_DOUBLE, P_BAD };
PType ptype(const string& t) {
  if (t == "char" || t == "int8") return P_CHAR;
  if (t == "uchar" || t == "uint8") return P_UCHAR;
  if (t == "short" || t == "int16") return P_SHORT;
  if (t == "ushort" || t == "uint16") return P_USHORT;
  if (t == "int" || t == "int32") return P_INT;
  if (t == "uint" || t == "uint32") return P_UINT;
  if (t == "float" || t == "float32") return P_FLOAT;
  if (t == "double" || t == "float64") return P_DOUBLE;
  return P_BAD;
}
int psize(PType t) {
  switch (t) {
    case P_CHAR: case P_UCHAR: return 1;
    case P_SHORT: case P_USHORT: return 2;
    case P_INT: case P_UINT: case P_FLOAT: return 4;
    case P_DOUBLE: return 8;
    default: return 0;
  }
}
struct Prop {
  string name;
  PType type = P_BAD;
  bool list = false;
  PType count_type = P_BAD;
};
struct Element {
  string name;
  uint64_t count = 0;
  vector<Prop> props;
};

struct Reader {
  FILE* f;
  int format;  // 0 ascii, 1 binary little endian, 2 binary big endian
  bool ok = true;
  double read(PType t) {
    if (format == 0) {
      double v = 0;
      if (fscanf(f, "%lf", &v) != 1) ok = false;
      return v;
    }
    uint8_t b[8];
    const int n = psize(t);
    if (fread(b, 1, n, f) != (size_t)n) {
      ok = false;
      return 0;
    }
    if (format == 2) std::reverse(b, b + n);
    switch (t) {
      case P_CHAR: return (double)(int8_t)b[0];
      case P_UCHAR: return (double)b[0];
      case P_SHORT: { int16_t v; memcpy(&v, b, 2); return v; }
      case P_USHORT: { uint16_t v; memcpy(&v, b, 2); return v; }
      case P_INT: { int32_t v; memcpy(&v, b, 4); return v; }
      case P_UINT: { uint32_t v; memcpy(&v, b, 4); return v; }
      case P_FLOAT: { float v; memcpy(&v, b, 4); return v; }
      case P_DOUBLE: { double v; memcpy(&v, b, 8); return v; }
      default: ok = false; return 0;
    }
  }
};

// float value of a colour channel: integer types are scaled to [0, 1] by their maximum
float colour_value(double v, PType t) {
  if (t == P_UCHAR) return (float)(v / 255.0);
  if (t == P_USHORT) return (float)(v / 65535.0);
  return (float)v;
}

bool read_ply(const string& path, Mesh& m, string& err) {
  FILE* f = fopen(path.c_str(), "rb");
  if (!f) {
    err = "cannot open " + path;
    return false;
  }
  char line[4096];
  vector<Element> elems;
  int format = -1;
  if (!fgets(line, sizeof line, f) || strncmp(line, "ply", 3) != 0) {
    fclose(f);
    err = "not a PLY file";
    return false;
  }
  while (fgets(line, sizeof line, f)) {
    istringstream ss(line);
    string w;
    ss >> w;
    if (w == "format") {
      string fmt;
      ss >> fmt;
      format = fmt == "ascii" ? 0 : fmt == "binary_little_endian" ? 1
             : fmt == "binary_big_endian" ? 2 : -1;
    } else if (w == "element") {
      Element e;
      ss >> e.name >> e.count;
      elems.push_back(e);
    } else if (w == "property" && !elems.empty()) {
      Prop p;
      string t;
      ss >> t;
      if (t == "list") {
        string ct, it;
        ss >> ct >> it >> p.name;
        p.list = true;
        p.count_type = ptype(ct);
        p.type = ptype(it);
      } else {
        p.type = ptype(t);
        ss >> p.name;
      }
      if (p.type == P_BAD || (p.list && p.count_type == P_BAD)) {
        fclose(f);
        err = "unsupported PLY property type";
        return false;
      }
      elems.back().props.push_back(p);
    } else if (w == "end_header") {
      break;
    }
  }
  if (format < 0) {
    fclose(f);
    err = "unsupported PLY format";
    return false;
  }
  Reader r{f, format};
  bool have_normals = false, have_colors = false;
  for (const Element& e : elems) {
    if (e.name == "vertex") {
      int ix = -1, iy = -1, iz = -1, inx = -1, iny = -1, inz = -1, ir = -1, ig = -1, ib = -1;
      for (size_t k = 0; k < e.props.size(); k++) {
        const string& n = e.props[k].name;
        if (n == "x") ix = (int)k;
        else if (n == "y") iy = (int)k;
        else if (n == "z") iz = (int)k;
        else if (n == "nx") inx = (int)k;
        else if (n == "ny") iny = (int)k;
        else if (n == "nz") inz = (int)k;
        else if (n == "red" || n == "diffuse_red" || n == "r") ir = (int)k;
        else if (n == "green" || n == "diffuse_green" || n == "g") ig = (int)k;
        else if (n == "blue" || n == "diffuse_blue" || n == "b") ib = (int)k;
      }
      have_normals = inx >= 0 && iny >= 0 && inz >= 0;
      have_colors = ir >= 0 && ig >= 0 && ib >= 0;
      m.vertices.resize(e.count);
      if (have_normals) m.normals.resize(e.count);
      if (have_colors) m.colors.resize(e.count);
      vector<double> vals(e.props.size());
      for (uint64_t i = 0; i < e.count; i++) {
        for (size_t k = 0; k < e.props.size(); k++) {
          const Prop& p = e.props[k];
          if (p.list) {
            const uint64_t n = (uint64_t)r.read(p.count_type);
            for (uint64_t q = 0; q < n; q++) r.read(p.type);
            vals[k] = 0;
          } else {
            vals[k] = r.read(p.type);
          }
        }
        if (ix >= 0) m.vertices[i].x = (float)vals[ix];
        if (iy >= 0) m.vertices[i].y = (float)vals[iy];
        if (iz >= 0) m.vertices[i].z = (float)vals[iz];
        if (have_normals)
          m.normals[i] = {(float)vals[inx], (float)vals[iny], (float)vals[inz]};
        if (have_colors)
          m.colors[i] = {colour_value(vals[ir], e.props[ir].type),
                         colour_value(vals[ig], e.props[ig].type),
                         colour_value(vals[ib], e.props[ib].type)};
      }
    } else if (e.name == "face") {
      for (uint64_t i = 0; i < e.count; i++) {
        for (const Prop& p : e.props) {
          if (!p.list) {
            r.read(p.type);
            continue;
          }
          const uint64_t n = (uint64_t)r.read(p.count_type);
          vector<uint32_t> idx(n);
          for (uint64_t q = 0; q < n; q++) idx[q] = (uint32_t)r.read(p.type);
          if (p.name != "vertex_indices" && p.name != "vertex_index") continue;
          for (uint64_t q = 1; q + 1 < n; q++) {  // triangle fan
            m.faces.push_back(idx[0]);
            m.faces.push_back(idx[q]);
            m.faces.push_back(idx[q + 1]);
          }
        }
      }
    } else {
      for (uint64_t i = 0; i < e.count; i++)
        for (const Prop& p : e.props) {
          if (p.list) {
            const uint64_t n = (uint64_t)r.read(p.count_type);
            for (uint64_t q = 0; q < n; q++) r.read(p.type);
          } else {
            r.read(p.type);
          }
        }
    }
    if (!r.ok) break;
  }
  fclose(f);
  if (!r.ok) {
    err = "PLY file ends early";
    return false;
  }
  return true;
}

// ------------------------------------------------------------------------------ OBJ
// v x y z [r g b]; vn nx ny nz; f with v, v/vt, v//vn or v/vt/vn corners (1-based, or
// negative = relative to the end). A corner's vn becomes its vertex's normal.
bool read_obj(const string& path, Mesh& m, string& err) {
  ifstream f(path);
  if (!f) {
    err = "cannot open " + path;
    return false;
  }
  vector<V3> vn;
  vector<int64_t> corner_vn;  // per vertex: index into vn, or -1
  bool have_colors = false;
  string line;
  while (getline(f, line)) {
    istringstream ss(line);
    string w;
    ss >> w;
    if (w == "v") {
      V3 p;
      ss >> p.x >> p.y >> p.z;
      m.vertices.push_back(p);
      V3 c;
      if (ss >> c.x >> c.y >> c.z) have_colors = true;
      m.colors.push_back(c);
      corner_vn.push_back(-1);
    } else if (w == "vn") {
      V3 n;
      ss >> n.x >> n.y >> n.z;
      vn.push_back(n);
    } else if (w == "f") {
      vector<int64_t> vi;
      string c;
      while (ss >> c) {
        const int64_t nv = (int64_t)m.vertices.size();
        int64_t a = atoll(c.c_str());
        a = a < 0 ? nv + a : a - 1;
        if (a < 0 || a >= nv) {
          err = "face index out of range";
          return false;
        }
        // the normal index is the third field (v/vt/vn or v//vn)
        const size_t s1 = c.find('/');
        if (s1 != string::npos) {
          const size_t s3 = c.find('/', s1 + 1);
          if (s3 != string::npos && s3 + 1 < c.size()) {
            int64_t n = atoll(c.c_str() + s3 + 1);
            n = n < 0 ? (int64_t)vn.size() + n : n - 1;
            if (n >= 0 && n < (int64_t)vn.size()) corner_vn[a] = n;
          }
        }
        vi.push_back(a);
      }
      for (size_t q = 1; q + 1 < vi.size(); q++) {
        m.faces.push_back((uint32_t)vi[0]);
        m.faces.push_back((uint32_t)vi[q]);
        m.faces.push_back((uint32_t)vi[q + 1]);
      }
    }
  }
  if (!have_colors) m.colors.clear();
  bool all = !vn.empty();
  for (int64_t c : corner_vn) all = all && c >= 0;
  if (all) {
    m.normals.resize(m.vertices.size());
    for (size_t i = 0; i < m.vertices.size(); i++) m.normals[i] = vn[corner_vn[i]];
  }
  return true;
}

#if TRI_PAYLOAD
// Vertex normals from faces with Max's weights, normalized.
void compute_vertex_normals(Mesh& m) {
  vector<V3> acc(m.vertices.size());
  for (size_t t = 0; t + 2 < m.faces.size(); t += 3) {
    const V3 p0 = m.vertices[m.faces[t]], p1 = m.vertices[m.faces[t + 1]],
             p2 = m.vertices[m.faces[t + 2]];
    const V3 a = sub(p0, p1), b = sub(p1, p2), c = sub(p2, p0);
    const float l2a = (a.x * a.x + a.y * a.y) + a.z * a.z;
    const float l2b = (b.x * b.x + b.y * b.y) + b.z * b.z;
    const float l2c = (c.x * c.x + c.y * c.y) + c.z * c.z;
    const V3 fn = cross(a, b);
    const float w[3] = {1.0f / (l2a * l2c), 1.0f / (l2b * l2a), 1.0f / (l2c * l2b)};
    for (int k = 0; k < 3; k++) {
      V3& n = acc[m.faces[t + k]];
      n.x = n.x + fn.x * w[k];
      n.y = n.y + fn.y * w[k];
      n.z = n.z + fn.z * w[k];
    }
  }
  m.normals.resize(m.vertices.size());
  for (size_t i = 0; i < acc.size(); i++) m.normals[i] = normalize(acc[i]);
}
#endif

}  // namespace

int main(int argc, char* argv[]) {
  printInfo();
  if (argc < 3) {
    printInvalid();
    return 0;
  }
  for (int i = 1; i < argc; i++) {
    const string a = argv[i];
    if (a == "-f" && i + 1 < argc) {
      filename = argv[++i];
    } else if (a == "-r") {
      recompute_normals = true;
    } else if (a == "-h") {
      printHelp();
      return 0;
    } else {
      printInvalid();
      return 0;
    }
  }
  cout << "  filename: " << filename << endl;
  cout << "  recompute normals: " << recompute_normals << endl;

  Mesh m;
  string err;
  string ext = filename.substr(filename.find_last_of('.') + 1);
  for (char& ch : ext) ch = (char)tolower(ch);
  const bool ok = ext == "ply" ? read_ply(filename, m, err)
                : ext == "obj" ? read_obj(filename, m, err)
                               : (err = "unsupported input format ." + ext, false);
  if (!ok) {
    cout << "Error: " << err << endl;
    return 1;
  }
  for (uint32_t v : m.faces)
    if (v >= m.vertices.size()) {
      cout << "Error: face index out of range" << endl;
      return 1;
    }
#if TRI_PAYLOAD
  if (m.normals.empty() && !recompute_normals) compute_vertex_normals(m);
#endif

  // bbox over all vertices, padded to a cube (createMeshBBCube, tri_convert_util.h:20-33)
  V3 mn, mx;
  if (!m.vertices.empty()) mn = mx = m.vertices[0];
  for (const V3& v : m.vertices) {
    mn.x = std::min(mn.x, v.x), mn.y = std::min(mn.y, v.y), mn.z = std::min(mn.z, v.z);
    mx.x = std::max(mx.x, v.x), mx.y = std::max(mx.y, v.y), mx.z = std::max(mx.z, v.z);
  }
  float lo[3] = {mn.x, mn.y, mn.z}, hi[3] = {mx.x, mx.y, mx.z};
  float len[3] = {hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]};
  const float maxlen = std::max(std::max(len[0], len[1]), len[2]);
  for (int k = 0; k < 3; k++) {
    const float delta = maxlen - len[k];
    if (delta != 0) {
      lo[k] = lo[k] - (delta / 2.0f);
      hi[k] = hi[k] + (delta / 2.0f);
    }
  }
  cout << "Moving mesh to origin ... ";
  for (V3& v : m.vertices) v = {v.x - lo[0], v.y - lo[1], v.z - lo[2]};
  cout << "done." << endl;

  const string base = filename.substr(0, filename.find_last_of('.'));
  const string header_name = base + ".tri", data_name = base + ".tridata";
  FILE* out = fopen(data_name.c_str(), "wb");
  if (!out) {
    cout << "Error: cannot write " << data_name << endl;
    return 1;
  }
  cout << "Writing mesh triangles ... ";
  const size_t ntri = m.faces.size() / 3;
  vector<float> rec;
  rec.reserve(1 << 16);
  for (size_t i = 0; i < ntri; i++) {
    const uint32_t a = m.faces[3 * i], b = m.faces[3 * i + 1], c = m.faces[3 * i + 2];
    const V3 v0 = m.vertices[a], v1 = m.vertices[b], v2 = m.vertices[c];
    const float geo[9] = {v0.x, v0.y, v0.z, v1.x, v1.y, v1.z, v2.x, v2.y, v2.z};
    rec.insert(rec.end(), geo, geo + 9);
#if TRI_PAYLOAD
    V3 n;
    if (recompute_normals) {
      n = normalize(cross(sub(v0, v1), sub(v1, v2)));  // computeFaceNormal
    } else {  // getShadingFaceNormal
      const V3 n0 = m.normals[a], n1 = m.normals[b], n2 = m.normals[c];
      V3 t;
      t.x = ((n0.x + n1.x) + n2.x) / 3.0f;
      t.y = ((n0.y + n1.y) + n2.y) / 3.0f;
      t.z = ((n0.z + n1.z) + n2.z) / 3.0f;
      n = normalize(t);
    }
    V3 c0, c1, c2;
    if (!m.colors.empty()) c0 = m.colors[a], c1 = m.colors[b], c2 = m.colors[c];
    const float pay[12] = {n.x, n.y, n.z, c0.x, c0.y, c0.z, c1.x, c1.y, c1.z, c2.x, c2.y, c2.z};
    rec.insert(rec.end(), pay, pay + 12);
#endif
    if (rec.size() >= (1 << 16) - 21) {
      fwrite(rec.data(), sizeof(float), rec.size(), out);
      rec.clear();
    }
  }
  fwrite(rec.data(), sizeof(float), rec.size(), out);
  fclose(out);
  cout << "done." << endl;

  // tri_tools.h:118-127 with the padded, untranslated bbox (tri_convert.cpp:139)
  cout << "Writing header to " << header_name << " ... " << endl;
  ofstream h(header_name);
  h << "#tri " << 1 << endl;
  h << "ntriangles " << ntri << endl;
  h << "geo_only " << (TRI_PAYLOAD ? 0 : 1) << endl;
  h << "bbox  " << lo[0] << " " << lo[1] << " " << lo[2] << " " << hi[0] << " " << hi[1] << " "
    << hi[2] << endl;
  h << "END" << endl;
  h.close();
  // TriInfo::print (tri_tools.h:23-30)
  cout << "  base_filename: " << endl
       << "  tri version: 1" << endl
       << "  geometry only: " << (TRI_PAYLOAD ? 0 : 1) << endl
       << "  n_triangles: " << ntri << endl
       << "  bbox min: " << lo[0] << " " << lo[1] << " " << lo[2] << endl
       << "  bbox max: " << hi[0] << " " << hi[1] << " " << hi[2] << endl;
  cout << "Done." << endl;
  return 0;
}
