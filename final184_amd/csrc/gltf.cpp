// gltf.cpp — glTF 2.0 reader for the voxelizer: a small JSON parser, the .glb container,
// base64 data URIs and external buffers, accessors, the node tree and mesh primitives.
// Host code only; see gltf.hpp for what is read and what it replaces.
#include "gltf.hpp"

#include <algorithm>
#include <charconv>
#include <cmath>
#include <cstring>
#include <fstream>
#include <sstream>

#include "../../include/vgi.h"

namespace vgi {

namespace {

// ---------------------------------------------------------------------------
// JSON (RFC 8259): objects keep their members in file order
// ---------------------------------------------------------------------------
struct Json {
  enum Type { kNull, kBool, kNum, kStr, kArr, kObj } type = kNull;
  bool boolean = false;
  double num = 0;
  std::string str;
  std::vector<Json> arr;
  std::vector<std::pair<std::string, Json>> obj;

  const Json* get(const char* key) const {
    if (type != kObj) return nullptr;
    for (const auto& kv : obj)
      if (kv.first == key) return &kv.second;
    return nullptr;
  }
  size_t size() const { return type == kArr ? arr.size() : 0; }
};

class JsonParser {
public:
  JsonParser(const char* p, size_t n) : start_(p), p_(p), end_(p + n) {}

  bool parse(Json& out, std::string& err) {
    if (!value(out, 0)) {
      err = "JSON: " + err_ + " at byte " + std::to_string(pos_ - start_);
      return false;
    }
    ws();
    if (p_ != end_) {
      err = "JSON: trailing characters";
      return false;
    }
    return true;
  }

private:
  static constexpr int kMaxDepth = 128;
  const char* start_;
  const char* p_;
  const char* end_;
  const char* pos_ = nullptr;
  std::string err_;

  bool fail(const char* m) {
    err_ = m;
    pos_ = p_;
    return false;
  }
  void ws() {
    while (p_ < end_ && (*p_ == ' ' || *p_ == '\t' || *p_ == '\n' || *p_ == '\r')) p_++;
  }
  bool lit(const char* s) {
    const size_t n = std::strlen(s);
    if ((size_t)(end_ - p_) < n || std::memcmp(p_, s, n) != 0) return false;
    p_ += n;
    return true;
  }
  static void utf8(uint32_t c, std::string& s) {
    if (c < 0x80) {
      s += (char)c;
    } else if (c < 0x800) {
      s += (char)(0xC0 | (c >> 6));
      s += (char)(0x80 | (c & 0x3F));
    } else if (c < 0x10000) {
      s += (char)(0xE0 | (c >> 12));
      s += (char)(0x80 | ((c >> 6) & 0x3F));
      s += (char)(0x80 | (c & 0x3F));
    } else {
      s += (char)(0xF0 | (c >> 18));
      s += (char)(0x80 | ((c >> 12) & 0x3F));
      s += (char)(0x80 | ((c >> 6) & 0x3F));
      s += (char)(0x80 | (c & 0x3F));
    }
  }
  bool hex4(uint32_t& v) {
    if (end_ - p_ < 4) return false;
    v = 0;
    for (int i = 0; i < 4; i++) {
      const char c = *p_++;
      v <<= 4;
      if (c >= '0' && c <= '9') v |= (uint32_t)(c - '0');
      else if (c >= 'a' && c <= 'f') v |= (uint32_t)(c - 'a' + 10);
      else if (c >= 'A' && c <= 'F') v |= (uint32_t)(c - 'A' + 10);
      else return false;
    }
    return true;
  }
  bool string(std::string& s) {
    if (p_ >= end_ || *p_ != '"') return fail("expected a string");
    p_++;
    s.clear();
    while (p_ < end_ && *p_ != '"') {
      const unsigned char c = (unsigned char)*p_;
      if (c < 0x20) return fail("control character in a string");
      if (c != '\\') {
        s += (char)c;
        p_++;
        continue;
      }
      if (++p_ >= end_) return fail("unterminated escape");
      const char e = *p_++;
      switch (e) {
        case '"': s += '"'; break;
        case '\\': s += '\\'; break;
        case '/': s += '/'; break;
        case 'b': s += '\b'; break;
        case 'f': s += '\f'; break;
        case 'n': s += '\n'; break;
        case 'r': s += '\r'; break;
        case 't': s += '\t'; break;
        case 'u': {
          uint32_t c1;
          if (!hex4(c1)) return fail("bad \\u escape");
          if (c1 >= 0xD800 && c1 < 0xDC00) {  // surrogate pair
            uint32_t c2;
            if (!lit("\\u") || !hex4(c2) || c2 < 0xDC00 || c2 >= 0xE000) return fail("bad surrogate pair");
            c1 = 0x10000 + ((c1 - 0xD800) << 10) + (c2 - 0xDC00);
          }
          utf8(c1, s);
          break;
        }
        default: return fail("bad escape");
      }
    }
    if (p_ >= end_) return fail("unterminated string");
    p_++;
    return true;
  }
  bool number(double& v) {
    const char* b = p_;
    if (p_ < end_ && *p_ == '-') p_++;
    while (p_ < end_ && ((*p_ >= '0' && *p_ <= '9') || *p_ == '.' || *p_ == 'e' || *p_ == 'E' || *p_ == '+' ||
                         *p_ == '-'))
      p_++;
    const auto r = std::from_chars(b, p_, v);  // locale-independent
    if (r.ec != std::errc() || r.ptr != p_) return fail("bad number");
    return true;
  }
  bool value(Json& v, int depth) {
    if (depth > kMaxDepth) return fail("nesting too deep");
    ws();
    if (p_ >= end_) return fail("unexpected end");
    const char c = *p_;
    if (c == '{') {
      p_++;
      v.type = Json::kObj;
      ws();
      if (p_ < end_ && *p_ == '}') {
        p_++;
        return true;
      }
      for (;;) {
        ws();
        std::pair<std::string, Json> kv;
        if (!string(kv.first)) return false;
        ws();
        if (p_ >= end_ || *p_ != ':') return fail("expected ':'");
        p_++;
        if (!value(kv.second, depth + 1)) return false;
        v.obj.push_back(std::move(kv));
        ws();
        if (p_ < end_ && *p_ == ',') {
          p_++;
          continue;
        }
        if (p_ < end_ && *p_ == '}') {
          p_++;
          return true;
        }
        return fail("expected ',' or '}'");
      }
    }
    if (c == '[') {
      p_++;
      v.type = Json::kArr;
      ws();
      if (p_ < end_ && *p_ == ']') {
        p_++;
        return true;
      }
      for (;;) {
        v.arr.emplace_back();
        if (!value(v.arr.back(), depth + 1)) return false;
        ws();
        if (p_ < end_ && *p_ == ',') {
          p_++;
          continue;
        }
        if (p_ < end_ && *p_ == ']') {
          p_++;
          return true;
        }
        return fail("expected ',' or ']'");
      }
    }
    if (c == '"') {
      v.type = Json::kStr;
      return string(v.str);
    }
    if (lit("true")) {
      v.type = Json::kBool;
      v.boolean = true;
      return true;
    }
    if (lit("false")) {
      v.type = Json::kBool;
      return true;
    }
    if (lit("null")) return true;
    v.type = Json::kNum;
    return number(v.num);
  }
};

// ---------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------
struct Fail {
  int code;
  std::string msg;
};

[[noreturn]] void raise(int code, const std::string& msg) { throw Fail{code, msg}; }

int64_t get_int(const Json& o, const char* key, int64_t def, bool required = false) {
  const Json* v = o.get(key);
  if (!v) {
    if (required) raise(VGI_E_INVALID, std::string("missing \"") + key + "\"");
    return def;
  }
  if (v->type != Json::kNum || v->num != std::floor(v->num) || std::fabs(v->num) > 9.0e15)
    raise(VGI_E_INVALID, std::string("\"") + key + "\" is not an integer");
  return (int64_t)v->num;
}

const Json& array_at(const Json* arr, int64_t i, const char* what) {
  if (!arr || arr->type != Json::kArr || i < 0 || (size_t)i >= arr->arr.size())
    raise(VGI_E_INVALID, std::string(what) + " index " + std::to_string(i) + " out of range");
  return arr->arr[(size_t)i];
}

bool base64(const std::string& in, size_t from, std::vector<uint8_t>& out) {
  auto val = [](char c) -> int {
    if (c >= 'A' && c <= 'Z') return c - 'A';
    if (c >= 'a' && c <= 'z') return c - 'a' + 26;
    if (c >= '0' && c <= '9') return c - '0' + 52;
    if (c == '+' || c == '-') return 62;
    if (c == '/' || c == '_') return 63;
    return -1;
  };
  uint32_t acc = 0;
  int bits = 0;
  for (size_t i = from; i < in.size(); i++) {
    const char c = in[i];
    if (c == '=') break;
    if (c == '\n' || c == '\r' || c == ' ') continue;
    const int v = val(c);
    if (v < 0) return false;
    acc = (acc << 6) | (uint32_t)v;
    bits += 6;
    if (bits >= 8) {
      bits -= 8;
      out.push_back((uint8_t)(acc >> bits));
    }
  }
  return true;
}

std::string percent_decode(const std::string& s) {
  std::string o;
  for (size_t i = 0; i < s.size(); i++) {
    if (s[i] == '%' && i + 2 < s.size() && std::isxdigit((unsigned char)s[i + 1]) &&
        std::isxdigit((unsigned char)s[i + 2])) {
      o += (char)std::stoi(s.substr(i + 1, 2), nullptr, 16);
      i += 2;
    } else {
      o += s[i];
    }
  }
  return o;
}

std::vector<uint8_t> read_file(const std::string& path) {
  std::ifstream f(path, std::ios::binary);
  if (!f) raise(VGI_E_INVALID, "cannot open " + path);
  std::ostringstream ss;
  ss << f.rdbuf();
  const std::string s = ss.str();
  return std::vector<uint8_t>(s.begin(), s.end());
}

uint32_t rd_u32(const uint8_t* p) {
  uint32_t v;
  std::memcpy(&v, p, 4);  // glTF is little-endian, as is the host
  return v;
}

// 4x4 column-major (GLSL / glTF) product r = a * b, in double
void mat_mul(const double* a, const double* b, double* r) {
  for (int c = 0; c < 4; c++)
    for (int i = 0; i < 4; i++) {
      double s = 0;
      for (int k = 0; k < 4; k++) s += a[k * 4 + i] * b[c * 4 + k];
      r[c * 4 + i] = s;
    }
}

// local transform of a node: its matrix, or T * R * S (glTF 2.0 spec 3.5.2)
void node_local(const Json& n, double* m) {
  if (const Json* mt = n.get("matrix")) {
    if (mt->type != Json::kArr || mt->arr.size() != 16) raise(VGI_E_INVALID, "node matrix is not 16 numbers");
    for (int i = 0; i < 16; i++) {
      if (mt->arr[i].type != Json::kNum) raise(VGI_E_INVALID, "node matrix is not 16 numbers");
      m[i] = mt->arr[i].num;
    }
    return;
  }
  double t[3] = {0, 0, 0}, q[4] = {0, 0, 0, 1}, s[3] = {1, 1, 1};
  const char* keys[3] = {"translation", "rotation", "scale"};
  double* dsts[3] = {t, q, s};
  const size_t lens[3] = {3, 4, 3};
  for (int k = 0; k < 3; k++) {
    const Json* a = n.get(keys[k]);
    if (!a) continue;
    if (a->type != Json::kArr || a->arr.size() != lens[k]) raise(VGI_E_INVALID, std::string("bad node ") + keys[k]);
    for (size_t i = 0; i < lens[k]; i++) {
      if (a->arr[i].type != Json::kNum) raise(VGI_E_INVALID, std::string("bad node ") + keys[k]);
      dsts[k][i] = a->arr[i].num;
    }
  }
  const double x = q[0], y = q[1], z = q[2], w = q[3];
  const double r[9] = {1 - 2 * (y * y + z * z), 2 * (x * y + z * w), 2 * (x * z - y * w),   // column 0
                       2 * (x * y - z * w), 1 - 2 * (x * x + z * z), 2 * (y * z + x * w),   // column 1
                       2 * (x * z + y * w), 2 * (y * z - x * w), 1 - 2 * (x * x + y * y)};  // column 2
  for (int c = 0; c < 3; c++) {
    for (int i = 0; i < 3; i++) m[c * 4 + i] = r[c * 3 + i] * s[c];
    m[c * 4 + 3] = 0;
  }
  m[12] = t[0];
  m[13] = t[1];
  m[14] = t[2];
  m[15] = 1;
}

class Reader {
public:
  Reader(const Json& root, std::vector<std::vector<uint8_t>>& buffers, uint32_t albedo_mode, GltfScene& out)
      : root_(root), buffers_(buffers), mode_(albedo_mode), out_(out) {}

  void run() {
    const Json* asset = root_.get("asset");
    const Json* ver = asset ? asset->get("version") : nullptr;
    if (!ver || ver->type != Json::kStr || ver->str.compare(0, 2, "2.") != 0)
      raise(VGI_E_INVALID, "not a glTF 2.x asset");
    if (const Json* req = root_.get("extensionsRequired"); req && req->size())
      raise(VGI_E_INVALID, "required extension " + req->arr[0].str + " is not supported");
    const Json* scenes = root_.get("scenes");
    if (!scenes || !scenes->size()) return;  // nothing to place
    const int64_t si = get_int(root_, "scene", 0);
    const Json& scene = array_at(scenes, si, "scene");
    const Json* nodes = scene.get("nodes");
    if (!nodes) return;
    const double I[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    for (const Json& n : nodes->arr) {
      if (n.type != Json::kNum) raise(VGI_E_INVALID, "scene node is not an index");
      visit((int64_t)n.num, I, 0);
    }
  }

private:
  const Json& root_;
  std::vector<std::vector<uint8_t>>& buffers_;
  uint32_t mode_;
  GltfScene& out_;
  std::vector<std::pair<int64_t, std::vector<uint32_t>>> mesh_cache_;  // mesh -> geometry ids (-1: skipped)

  // node tree, depth first, children in order (glTFSceneImporter.cpp:333-372)
  void visit(int64_t ni, const double* parent, int depth) {
    if (depth > 64) raise(VGI_E_INVALID, "node tree deeper than 64 (cycle?)");
    const Json& n = array_at(root_.get("nodes"), ni, "node");
    double local[16], world[16];
    node_local(n, local);
    mat_mul(parent, local, world);
    const int64_t mi = get_int(n, "mesh", -1);
    if (mi >= 0) {
      const std::vector<uint32_t>& geo = mesh(mi);
      const Json& m = array_at(root_.get("meshes"), mi, "mesh");
      const Json* prims = m.get("primitives");
      for (size_t p = 0; p < geo.size(); p++) {
        if (geo[p] == UINT32_MAX) continue;
        const Json& pr = prims->arr[p];
        const int64_t mat = get_int(pr, "material", -1);
        if (mode_ == 0 && mat < 0) continue;  // the reference voxelizer draws only primitives with a material
        GltfInstance inst;
        inst.geometry = geo[p];
        inst.node = (int32_t)ni;
        inst.material = (int32_t)mat;
        for (int r = 0; r < 3; r++)
          for (int c = 0; c < 4; c++) inst.model[r * 4 + c] = (float)world[c * 4 + r];
        base_color(mat, inst.base_color);
        texture(mat, inst);
        out_.instances.push_back(inst);
      }
    }
    if (const Json* ch = n.get("children"))
      for (const Json& c : ch->arr) {
        if (c.type != Json::kNum) raise(VGI_E_INVALID, "child is not an index");
        visit((int64_t)c.num, world, depth + 1);
      }
  }

  void base_color(int64_t mat, uint8_t* rgba) {
    double f[4] = {1, 1, 1, 1};
    if (mode_ == 1 && mat >= 0) {
      const Json& m = array_at(root_.get("materials"), mat, "material");
      const Json* pbr = m.get("pbrMetallicRoughness");
      const Json* bc = pbr ? pbr->get("baseColorFactor") : nullptr;
      if (bc) {
        if (bc->type != Json::kArr || bc->arr.size() != 4) raise(VGI_E_INVALID, "bad baseColorFactor");
        for (int k = 0; k < 4; k++) f[k] = bc->arr[k].num;
      }
    }
    for (int k = 0; k < 4; k++) rgba[k] = (uint8_t)std::lround(std::fmin(std::fmax(f[k], 0.0), 1.0) * 255.0);
  }

  // The material's base colour texture and the factor it is multiplied by. The reference's
  // importer (GetMaterial, glTFSceneImporter.cpp:197-201) binds GetImage(texture index):
  // it indexes the images array with the texture's index (GetImage, :110-131), not with
  // the texture's "source", and leaves the factor white. Mode 0 does the same; mode 1
  // follows the texture to its source image and uses baseColorFactor. Either way the
  // texture is sampled with TEXCOORD_0 (the reference's shaders have no other set).
  void texture(int64_t mat, GltfInstance& inst) {
    for (int k = 0; k < 4; k++) inst.factor[k] = 1.0f;
    inst.image = -1;
    if (mat < 0) return;
    const Json& m = array_at(root_.get("materials"), mat, "material");
    const Json* pbr = m.get("pbrMetallicRoughness");
    if (!pbr) return;
    if (mode_ == 1)
      if (const Json* bc = pbr->get("baseColorFactor"))
        for (int k = 0; k < 4 && k < (int)bc->size(); k++) inst.factor[k] = (float)bc->arr[k].num;
    const Json* bt = pbr->get("baseColorTexture");
    if (!bt) return;
    const int64_t ti = get_int(*bt, "index", -1, true);
    const Json& t = array_at(root_.get("textures"), ti, "texture");
    const int64_t src = mode_ == 0 ? ti : get_int(t, "source", -1);
    if (src < 0) return;
    array_at(root_.get("images"), src, "image");  // in range
    inst.image = (int32_t)src;
  }

  // the bytes an accessor covers: element e starts at base + e * stride
  struct View {
    const uint8_t* base = nullptr;  // nullptr: no bufferView (all zeros)
    size_t stride = 0;
    size_t count = 0;
    int comp = 0;
  };

  View accessor(int64_t ai, const char* type, size_t ncomp) {
    const Json& a = array_at(root_.get("accessors"), ai, "accessor");
    if (a.get("sparse")) raise(VGI_E_INVALID, "sparse accessors are not supported");
    const Json* t = a.get("type");
    if (!t || t->type != Json::kStr || t->str != type)
      raise(VGI_E_INVALID, "accessor " + std::to_string(ai) + " is not " + type);
    View v;
    v.comp = (int)get_int(a, "componentType", 0, true);
    const int64_t count = get_int(a, "count", 0, true);
    if (count < 0 || count > (int64_t)UINT32_MAX) raise(VGI_E_RANGE, "accessor count out of range");
    v.count = (size_t)count;
    size_t csize = 0;
    switch (v.comp) {
      case 5120: case 5121: csize = 1; break;
      case 5122: case 5123: csize = 2; break;
      case 5125: case 5126: csize = 4; break;
      default: raise(VGI_E_INVALID, "bad componentType " + std::to_string(v.comp));
    }
    const size_t esize = csize * ncomp;
    const int64_t bvi = get_int(a, "bufferView", -1);
    if (bvi < 0) {
      v.stride = esize;
      return v;
    }
    const Json& bv = array_at(root_.get("bufferViews"), bvi, "bufferView");
    const int64_t bi = get_int(bv, "buffer", 0, true);
    if (bi < 0 || (size_t)bi >= buffers_.size()) raise(VGI_E_INVALID, "bufferView buffer out of range");
    const std::vector<uint8_t>& buf = buffers_[(size_t)bi];
    const int64_t bo = get_int(bv, "byteOffset", 0), bl = get_int(bv, "byteLength", 0, true);
    const int64_t st = get_int(bv, "byteStride", 0), ao = get_int(a, "byteOffset", 0);
    if (bo < 0 || bl < 0 || ao < 0 || st < 0 || (uint64_t)bo + (uint64_t)bl > buf.size())
      raise(VGI_E_INVALID, "bufferView outside its buffer");
    v.stride = st ? (size_t)st : esize;
    if (v.stride < esize) raise(VGI_E_INVALID, "byteStride smaller than the element");
    if (v.count && (uint64_t)ao + (uint64_t)v.stride * (v.count - 1) + esize > (uint64_t)bl)
      raise(VGI_E_INVALID, "accessor " + std::to_string(ai) + " reads past its bufferView");
    v.base = buf.data() + bo + ao;
    return v;
  }

  // geometry of every primitive of a mesh (glTFSceneImporter.cpp:210-326), made once
  const std::vector<uint32_t>& mesh(int64_t mi) {
    for (const auto& e : mesh_cache_)
      if (e.first == mi) return e.second;
    const Json& m = array_at(root_.get("meshes"), mi, "mesh");
    const Json* prims = m.get("primitives");
    if (!prims || prims->type != Json::kArr) raise(VGI_E_INVALID, "mesh without primitives");
    std::vector<uint32_t> ids;
    for (size_t p = 0; p < prims->arr.size(); p++) {
      const Json& pr = prims->arr[p];
      const int64_t mode = get_int(pr, "mode", 4);
      if (mode < 0 || mode > 6) raise(VGI_E_INVALID, "bad primitive mode");
      if (mode < 4) {  // points and lines: nothing to voxelize
        out_.skipped_primitives++;
        ids.push_back(UINT32_MAX);
        continue;
      }
      const Json* attrs = pr.get("attributes");
      const int64_t pa = attrs ? get_int(*attrs, "POSITION", -1) : -1;
      if (pa < 0) raise(VGI_E_INVALID, "primitive without POSITION");
      const View pv = accessor(pa, "VEC3", 3);
      if (pv.comp != 5126) raise(VGI_E_INVALID, "POSITION is not float");
      GltfGeometry g;
      g.mesh = (int32_t)mi;
      g.primitive = (int32_t)p;
      g.positions.resize(pv.count * 3);
      for (size_t e = 0; e < pv.count; e++)
        if (pv.base) std::memcpy(&g.positions[e * 3], pv.base + e * pv.stride, 12);
        else g.positions[e * 3] = g.positions[e * 3 + 1] = g.positions[e * 3 + 2] = 0.f;
      // TEXCOORD_0: float, or normalized unsigned byte / short (c / 255, c / 65535)
      const int64_t ta = get_int(*attrs, "TEXCOORD_0", -1);
      if (ta >= 0) {
        const View tv = accessor(ta, "VEC2", 2);
        if (tv.count != pv.count) raise(VGI_E_INVALID, "TEXCOORD_0 and POSITION counts differ");
        if (tv.comp != 5126 && tv.comp != 5121 && tv.comp != 5123)
          raise(VGI_E_INVALID, "TEXCOORD_0 is not float, unsigned byte or unsigned short");
        g.texcoords.assign(tv.count * 2, 0.f);
        for (size_t e = 0; e < tv.count && tv.base; e++) {
          const uint8_t* q = tv.base + e * tv.stride;
          for (int c = 0; c < 2; c++) {
            if (tv.comp == 5126) {
              std::memcpy(&g.texcoords[e * 2 + c], q + 4 * c, 4);
            } else if (tv.comp == 5121) {
              g.texcoords[e * 2 + c] = (float)q[c] / 255.0f;
            } else {
              uint16_t h;
              std::memcpy(&h, q + 2 * c, 2);
              g.texcoords[e * 2 + c] = (float)h / 65535.0f;
            }
          }
        }
      }
      // vertex order: the index accessor, or 0..count-1
      std::vector<uint32_t> order;
      const int64_t ia = get_int(pr, "indices", -1);
      if (ia >= 0) {
        const View iv = accessor(ia, "SCALAR", 1);
        order.resize(iv.count);
        for (size_t e = 0; e < iv.count; e++) {
          uint32_t x = 0;
          if (iv.base) {
            const uint8_t* q = iv.base + e * iv.stride;
            if (iv.comp == 5121) x = q[0];
            else if (iv.comp == 5123) { uint16_t h; std::memcpy(&h, q, 2); x = h; }
            else if (iv.comp == 5125) std::memcpy(&x, q, 4);
            else raise(VGI_E_INVALID, "index componentType must be unsigned");
          }
          if (x >= pv.count) raise(VGI_E_INVALID, "index " + std::to_string(x) + " >= vertex count");
          order[e] = x;
        }
      } else {
        order.resize(pv.count);
        for (size_t e = 0; e < pv.count; e++) order[e] = (uint32_t)e;
      }
      // primitive assembly (Vulkan: lists; strips (i, i+1+i%2, i+2-i%2); fans (i+1, i+2, 0))
      const size_t nv = order.size();
      if (mode == 4) {
        g.indices.assign(order.begin(), order.begin() + (nv / 3) * 3);
      } else if (nv >= 3) {
        g.indices.reserve((nv - 2) * 3);
        for (size_t i = 0; i + 2 < nv; i++) {
          if (mode == 5) {
            g.indices.push_back(order[i]);
            g.indices.push_back(order[i + 1 + (i & 1)]);
            g.indices.push_back(order[i + 2 - (i & 1)]);
          } else {
            g.indices.push_back(order[i + 1]);
            g.indices.push_back(order[i + 2]);
            g.indices.push_back(order[0]);
          }
        }
      }
      ids.push_back((uint32_t)out_.geometry.size());
      out_.geometry.push_back(std::move(g));
    }
    mesh_cache_.emplace_back(mi, std::move(ids));
    return mesh_cache_.back().second;
  }
};

}  // namespace

int gltf_load_image(const GltfScene& scene, GltfImage& img, std::string& err) {
  if (img.loaded) return VGI_OK;
  if (scene.base_dir.empty()) {
    err = "external image " + img.uri + " without a base directory";
    return VGI_E_INVALID;
  }
  try {
    img.data = read_file(scene.base_dir + "/" + img.uri);
  } catch (const Fail& f) {
    err = f.msg;
    return f.code;
  } catch (const std::bad_alloc&) {
    err = "out of host memory";
    return VGI_E_NOMEM;
  }
  img.loaded = true;
  return VGI_OK;
}

int gltf_parse(const uint8_t* data, size_t bytes, const std::string& base_dir, uint32_t albedo_mode,
               GltfScene& out, std::string& err) {
  out = GltfScene{};
  if (albedo_mode > 1) {
    err = "albedo_mode must be 0 (reference) or 1 (baseColorFactor)";
    return VGI_E_INVALID;
  }
  try {
    const char* json = (const char*)data;
    size_t json_len = bytes;
    std::vector<uint8_t> bin;
    bool glb = false;
    if (bytes >= 12 && rd_u32(data) == 0x46546C67u) {  // "glTF"
      glb = true;
      if (rd_u32(data + 4) != 2) raise(VGI_E_INVALID, "GLB version is not 2");
      const size_t total = rd_u32(data + 8);
      if (total > bytes || total < 20) raise(VGI_E_INVALID, "GLB length field exceeds the data");
      size_t off = 12;
      json = nullptr;
      while (off + 8 <= total) {
        const size_t len = rd_u32(data + off);
        const uint32_t type = rd_u32(data + off + 4);
        if (len > total - off - 8) raise(VGI_E_INVALID, "GLB chunk exceeds the file");
        if (type == 0x4E4F534Au && !json) {  // JSON
          json = (const char*)data + off + 8;
          json_len = len;
        } else if (type == 0x004E4942u && bin.empty()) {  // BIN
          bin.assign(data + off + 8, data + off + 8 + len);
        }
        off += 8 + ((len + 3) & ~(size_t)3);
      }
      if (!json) raise(VGI_E_INVALID, "GLB without a JSON chunk");
    }
    Json root;
    JsonParser parser(json, json_len);
    if (!parser.parse(root, err)) return VGI_E_INVALID;
    if (root.type != Json::kObj) raise(VGI_E_INVALID, "top level is not an object");
    std::vector<std::vector<uint8_t>> buffers;
    if (const Json* bl = root.get("buffers")) {
      for (size_t i = 0; i < bl->size(); i++) {
        const Json& b = bl->arr[i];
        const int64_t len = get_int(b, "byteLength", 0, true);
        const Json* uri = b.get("uri");
        std::vector<uint8_t> d;
        if (!uri) {
          if (!glb || i != 0) raise(VGI_E_INVALID, "buffer without uri outside a GLB");
          d = bin;
        } else if (uri->str.compare(0, 5, "data:") == 0) {
          const size_t comma = uri->str.find(',');
          if (comma == std::string::npos || uri->str.find(";base64") == std::string::npos ||
              uri->str.find(";base64") > comma || !base64(uri->str, comma + 1, d))
            raise(VGI_E_INVALID, "buffer " + std::to_string(i) + ": unsupported data URI");
        } else {
          const std::string u = percent_decode(uri->str);
          if (base_dir.empty()) raise(VGI_E_INVALID, "external buffer " + u + " without a base directory");
          if (u.find("://") != std::string::npos || (!u.empty() && u[0] == '/'))
            raise(VGI_E_INVALID, "buffer uri " + u + " is not a relative path");
          d = read_file(base_dir + "/" + u);
        }
        if (len < 0 || (size_t)len > d.size()) raise(VGI_E_INVALID, "buffer " + std::to_string(i) + " is shorter than byteLength");
        d.resize((size_t)len);
        buffers.push_back(std::move(d));
      }
    }
    Reader r(root, buffers, albedo_mode, out);
    r.run();
    // images as stored: a bufferView's bytes, a base64 data URI, or a relative file read
    // on request (the reference decodes them with stb_image inside tinygltf; decoding is
    // the caller's here)
    out.base_dir = base_dir;
    if (const Json* il = root.get("images")) {
      for (size_t i = 0; i < il->size(); i++) {
        const Json& im = il->arr[i];
        GltfImage g;
        if (const Json* mt = im.get("mimeType"))
          if (mt->type == Json::kStr) g.mime_type = mt->str;
        const Json* uri = im.get("uri");
        const int64_t bvi = get_int(im, "bufferView", -1);
        if (uri && uri->type == Json::kStr && uri->str.compare(0, 5, "data:") == 0) {
          const size_t comma = uri->str.find(',');
          const size_t b64 = uri->str.find(";base64");
          if (comma == std::string::npos || b64 == std::string::npos || b64 > comma ||
              !base64(uri->str, comma + 1, g.data))
            raise(VGI_E_INVALID, "image " + std::to_string(i) + ": unsupported data URI");
          if (g.mime_type.empty()) g.mime_type = uri->str.substr(5, std::min(b64, uri->str.find(';')) - 5);
          g.loaded = true;
        } else if (uri && uri->type == Json::kStr) {
          g.uri = percent_decode(uri->str);
          if (g.uri.find("://") != std::string::npos || (!g.uri.empty() && g.uri[0] == '/'))
            raise(VGI_E_INVALID, "image uri " + g.uri + " is not a relative path");
        } else if (bvi >= 0) {
          const Json& bv = array_at(root.get("bufferViews"), bvi, "bufferView");
          const int64_t bi = get_int(bv, "buffer", 0, true);
          if (bi < 0 || (size_t)bi >= buffers.size()) raise(VGI_E_INVALID, "bufferView buffer out of range");
          const int64_t bo = get_int(bv, "byteOffset", 0), bl = get_int(bv, "byteLength", 0, true);
          if (bo < 0 || bl < 0 || (uint64_t)bo + (uint64_t)bl > buffers[(size_t)bi].size())
            raise(VGI_E_INVALID, "bufferView outside its buffer");
          g.data.assign(buffers[(size_t)bi].begin() + bo, buffers[(size_t)bi].begin() + bo + bl);
          g.loaded = true;
        } else {
          raise(VGI_E_INVALID, "image " + std::to_string(i) + " has neither uri nor bufferView");
        }
        out.images.push_back(std::move(g));
      }
    }
  } catch (const Fail& f) {
    err = f.msg;
    out = GltfScene{};
    return f.code;
  } catch (const std::bad_alloc&) {
    err = "out of host memory";
    out = GltfScene{};
    return VGI_E_NOMEM;
  } catch (const std::exception& e) {
    err = std::string("malformed file: ") + e.what();
    out = GltfScene{};
    return VGI_E_INVALID;
  }
  return VGI_OK;
}

}  // namespace vgi
