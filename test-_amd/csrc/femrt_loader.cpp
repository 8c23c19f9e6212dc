// femrt_loader.cpp -- Nektar++ XML mesh + FLD field reader and writer for
// hexahedral meshes (the drop-in for load_xml_fld, xml_fld_loader.cpp:16-402).
//
// Reader: VERTEX / EDGE / FACE / ELEMENT (H as 6 face ids or 8 vertex ids) /
// CURVED (E and F point sets, PolyEvenlySpaced or GaussLobattoLegendre) /
// COMPOSITE / EXPANSIONS from the XML; every ELEMENTS block of the FLD
// (ASCII numbers or B64Z-LittleEndian, 32- or 64-bit).  Geometry per hex as
// the reference builds it (BuildHexGeometry, xml_fld_geometry_builder.cpp:
// 68-381): trilinear map on the GLL grid, curved edge points resampled to the
// GLL nodes and placed on the edges, curved faces resampled to the GLL grid
// and blended linearly along the face-normal coordinate, tensor modal fit.
// Differences from the reference, each fixed by default and replicable by a
// flag: the first vertex is parsed (the reference's scan matches the
// <VERTEX> tag itself and sets vertex 0 to the origin,
// xml_fld_geometry_parser.cpp:54-77; flag bit0), geometry modes come from the
// curved data alone (the reference raises them to the field modes,
// xml_fld_loader.cpp:245-254; flag bit1).  NUMMODESPERDIR="UNIORDER:a,b,c"
// is read as (a, b, c) (the reference's ParseModes reads 0 for the first
// entry, xml_fld_common.cpp:63-100); the FLD's ID lists place coefficients
// (the reference assigns them in expansion order from the first ELEMENTS
// block only, xml_fld_field_parser.cpp:36-51).
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>
#include <zlib.h>

#include <algorithm>
#include <array>
#include <cctype>
#include <charconv>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <memory>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "femrt_internal.hpp"

namespace femrt {

namespace {

struct P3 {
  double x = 0, y = 0, z = 0;
};

// A read-only view of a range of characters.
struct Text {
  const char* b = nullptr;
  const char* e = nullptr;
  bool empty() const { return b >= e; }
  std::string str() const { return std::string(b, e); }
};

// A file mapped read-only: the scanner reads it in place, page by page as it
// goes, with no copy of the text.
struct MappedFile {
  int fd = -1;
  void* p = nullptr;
  size_t n = 0;
  long long mtime_ns = 0;
  bool open(const char* path)
  {
    fd = ::open(path, O_RDONLY);
    if (fd < 0) return false;
    struct stat st;
    if (fstat(fd, &st) != 0) return false;
    n = (size_t)st.st_size;
    mtime_ns = (long long)st.st_mtim.tv_sec * 1000000000LL + st.st_mtim.tv_nsec;
    if (n == 0) return true;
    p = mmap(nullptr, n, PROT_READ, MAP_PRIVATE, fd, 0);
    if (p == MAP_FAILED) {
      p = nullptr;
      return false;
    }
    madvise(p, n, MADV_SEQUENTIAL);
    return true;
  }
  Text text() const { return Text{(const char*)p, (const char*)p + (p ? n : 0)}; }
  ~MappedFile()
  {
    if (p) munmap(p, n);
    if (fd >= 0) ::close(fd);
  }
};

inline char up(char c) { return (c >= 'a' && c <= 'z') ? (char)(c - 32) : c; }
inline bool is_space(char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\r'; }

// Case-insensitive match of the upper-case literal U at p (p < e).
inline bool imatch(const char* p, const char* e, const char* U)
{
  for (; *U; ++U, ++p)
    if (p >= e || up(*p) != *U) return false;
  return true;
}

std::string upper(std::string s)
{
  for (auto& c : s) c = up(c);
  return s;
}

// The next opening tag <T ...> (T upper case, matched whole and
// case-insensitive) at or after p; comments <!-- ... --> are skipped.
const char* find_open(const char* p, const char* e, const char* T)
{
  const size_t tl = std::strlen(T);
  while (p < e) {
    p = (const char*)std::memchr(p, '<', (size_t)(e - p));
    if (!p) return nullptr;
    if (imatch(p, e, "<!--")) {
      const char* q = p + 4;
      while (q + 2 < e && !(q[0] == '-' && q[1] == '-' && q[2] == '>')) ++q;
      p = q + 3;
      continue;
    }
    if (imatch(p + 1, e, T)) {
      const char* q = p + 1 + tl;
      if (q < e && (is_space(*q) || *q == '>' || *q == '/')) return p;
    }
    ++p;
  }
  return nullptr;
}

// The closing tag </T> at or after p (case-insensitive).
const char* find_close(const char* p, const char* e, const char* T)
{
  while (p < e) {
    p = (const char*)std::memchr(p, '<', (size_t)(e - p));
    if (!p) return nullptr;
    if (p + 1 < e && p[1] == '/' && imatch(p + 2, e, T)) {
      const char* q = p + 2 + std::strlen(T);
      while (q < e && is_space(*q)) ++q;
      if (q < e && *q == '>') return p;
    }
    ++p;
  }
  return nullptr;
}

// Value of attribute key in an opening tag (case-insensitive key).
bool attr(Text tag, const char* key, std::string& out)
{
  const std::string K = upper(key);
  for (const char* p = tag.b; p < tag.e; ++p) {
    if (!imatch(p, tag.e, K.c_str())) continue;
    const bool left_ok = p > tag.b && (is_space(p[-1]) || p[-1] == '<');
    const char* q = p + K.size();
    while (q < tag.e && is_space(*q)) ++q;
    if (!left_ok || q >= tag.e || *q != '=') continue;
    ++q;
    while (q < tag.e && is_space(*q)) ++q;
    if (q >= tag.e || (*q != '"' && *q != '\'')) continue;
    const char quote = *q;
    const char* v = q + 1;
    const char* w = (const char*)std::memchr(v, quote, (size_t)(tag.e - v));
    if (!w) return false;
    out.assign(v, w);
    return true;
  }
  return false;
}

// Body of the section <name ...> ... </name>, empty when absent.
Text section(Text s, const char* name)
{
  const std::string N = upper(name);
  const char* p = find_open(s.b, s.e, N.c_str());
  if (!p) return {};
  const char* gt = (const char*)std::memchr(p, '>', (size_t)(s.e - p));
  if (!gt) return {};
  const char* c = find_close(gt + 1, s.e, N.c_str());
  if (!c) return {};
  return Text{gt + 1, c};
}

// Calls f(header, body) for every element <tag ...>body</tag> or <tag .../>
// (tag matched whole, case-insensitive), in order.
template <typename F>
void each(Text s, const char* tag, F f)
{
  const std::string T = upper(tag);
  const char* p = s.b;
  while ((p = find_open(p, s.e, T.c_str())) != nullptr) {
    const char* gt = (const char*)std::memchr(p, '>', (size_t)(s.e - p));
    if (!gt) return;
    const Text header{p, gt + 1};
    if (gt > p && gt[-1] == '/') {
      f(header, Text{});
      p = gt + 1;
      continue;
    }
    const char* c = find_close(gt + 1, s.e, T.c_str());
    if (!c) return;
    f(header, Text{gt + 1, c});
    p = c + 2 + T.size();
  }
}

// "0-3,7,9-10" -> 0 1 2 3 7 9 10
std::vector<int> ranges(const std::string& s)
{
  std::vector<int> out;
  std::string tok;
  std::stringstream ss(s);
  while (std::getline(ss, tok, ',')) {
    const size_t d = tok.find('-', tok.find_first_not_of(" \t\r\n"));
    if (d != std::string::npos && d > 0) {
      const int lo = std::atoi(tok.substr(0, d).c_str()), hi = std::atoi(tok.substr(d + 1).c_str());
      for (int v = lo; v <= hi; ++v) out.push_back(v);
    } else if (tok.find_first_of("0123456789") != std::string::npos) {
      out.push_back(std::atoi(tok.c_str()));
    }
  }
  return out;
}

// Mode counts of "5", "5,5,5" or "UNIORDER:5,5,5"
bool parse_modes(const std::string& s, int m[3])
{
  std::string v = s;
  const size_t c = v.find(':');
  if (c != std::string::npos) v = v.substr(c + 1);
  std::vector<int> n;
  std::string tok;
  std::stringstream ss(v);
  while (std::getline(ss, tok, ',')) n.push_back(std::atoi(tok.c_str()));
  if (n.size() == 1) n = {n[0], n[0], n[0]};
  if (n.size() != 3) return false;
  for (int k = 0; k < 3; ++k) m[k] = n[k];
  return true;
}

std::vector<uint8_t> b64_decode(Text in)
{
  static int lut[256];
  static bool init = false;
  if (!init) {
    for (int& v : lut) v = -1;
    const char* a = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/";
    for (int k = 0; k < 64; ++k) lut[(unsigned char)a[k]] = k;
    init = true;
  }
  std::vector<uint8_t> out;
  out.reserve((size_t)(in.e - in.b) * 3 / 4);
  unsigned acc = 0;
  int bits = 0;
  for (const char* p = in.b; p < in.e; ++p) {
    const unsigned char c = (unsigned char)*p;
    if (c == '=') break;
    const int v = lut[c];
    if (v < 0) continue;  // whitespace
    acc = (acc << 6) | (unsigned)v;
    bits += 6;
    if (bits >= 8) {
      bits -= 8;
      out.push_back((uint8_t)((acc >> bits) & 0xFF));
    }
  }
  return out;
}

bool inflate_all(const std::vector<uint8_t>& in, std::vector<uint8_t>& out)
{
  z_stream zs;
  std::memset(&zs, 0, sizeof(zs));
  if (inflateInit(&zs) != Z_OK) return false;
  zs.next_in = const_cast<Bytef*>(in.data());
  zs.avail_in = (uInt)in.size();
  out.clear();
  uint8_t buf[1 << 15];
  int rc;
  do {
    zs.next_out = buf;
    zs.avail_out = sizeof(buf);
    rc = inflate(&zs, Z_NO_FLUSH);
    if (rc != Z_OK && rc != Z_STREAM_END) {
      inflateEnd(&zs);
      return false;
    }
    out.insert(out.end(), buf, buf + (sizeof(buf) - zs.avail_out));
  } while (rc != Z_STREAM_END);
  inflateEnd(&zs);
  return true;
}

std::string b64_encode(const uint8_t* d, size_t n)
{
  const char* a = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/";
  std::string out;
  out.reserve((n + 2) / 3 * 4);
  for (size_t i = 0; i < n; i += 3) {
    const unsigned v = (unsigned)d[i] << 16 | (i + 1 < n ? (unsigned)d[i + 1] << 8 : 0u) |
                       (i + 2 < n ? (unsigned)d[i + 2] : 0u);
    out.push_back(a[(v >> 18) & 63]);
    out.push_back(a[(v >> 12) & 63]);
    out.push_back(i + 1 < n ? a[(v >> 6) & 63] : '=');
    out.push_back(i + 2 < n ? a[v & 63] : '=');
  }
  return out;
}

// Whitespace-separated numbers from the start of s (appended to v); *stop
// (may be NULL) receives where parsing stopped.
void parse_numbers(Text s, std::vector<double>& v, const char** stop = nullptr)
{
  const char* p = s.b;
  for (;;) {
    while (p < s.e && is_space(*p)) ++p;
    if (p >= s.e) break;
    const char* q = (*p == '+') ? p + 1 : p;
    double x;
    const auto r = std::from_chars(q, s.e, x);
    if (r.ec != std::errc()) break;
    v.push_back(x);
    p = r.ptr;
  }
  if (stop) *stop = p;
}

std::vector<double> numbers(Text s)
{
  std::vector<double> v;
  parse_numbers(s, v);
  return v;
}

// Hex vertex order from the 6 faces in Nektar order: vertex c is the one
// common to the three faces that meet at corner c (the format's convention;
// the reference derives it the same way, xml_fld_geometry_parser.cpp:227-236).
const int kCornerFaces[8][3] = {{0, 1, 4}, {0, 1, 2}, {0, 2, 3}, {0, 3, 4},
                                {1, 4, 5}, {1, 2, 5}, {2, 3, 5}, {3, 4, 5}};
// local faces: corners spanning (u, v) as (0,0), (max,0), (max,max), (0,max)
const int kFaceCorners[6][4] = {{0, 1, 2, 3}, {0, 1, 5, 4}, {1, 2, 6, 5},
                                {3, 2, 6, 7}, {0, 3, 7, 4}, {4, 5, 6, 7}};
// local edges: low and high corner along the edge's parameter, and the axis
const int kEdges[12][3] = {{0, 1, 0}, {1, 2, 1}, {3, 2, 0}, {0, 3, 1}, {0, 4, 2}, {1, 5, 2},
                           {2, 6, 2}, {3, 7, 2}, {4, 5, 0}, {5, 6, 1}, {7, 6, 0}, {4, 7, 1}};
// reference-space position of each corner
const int kCornerRst[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                              {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};

struct Curved {
  int npts = 0;
  bool gll = false;
  std::vector<P3> pts;
};

std::vector<double> nodes_of(int n, bool gll)
{
  if (gll) return gll_nodes(n);
  std::vector<double> x(n);
  for (int i = 0; i < n; ++i) x[i] = n == 1 ? 0.0 : -1.0 + 2.0 * i / (n - 1);
  return x;
}

double dist2(const P3& a, const P3& b)
{
  return (a.x - b.x) * (a.x - b.x) + (a.y - b.y) * (a.y - b.y) + (a.z - b.z) * (a.z - b.z);
}

}  // namespace

namespace {

struct Hex {
  int id;
  std::array<int, 8> v;
  std::array<int, 6> f;
  bool has_faces;
};

// Everything the reader takes from the two files, before the per-element
// output: kept between the size query and the fill call of the two-call
// protocol, so each file is scanned once.
struct Parsed {
  std::string xml_path, fld_path;
  int flags = 0;
  long long xml_size = -1, xml_mtime = 0, fld_size = -1, fld_mtime = 0;
  std::unordered_map<int, P3> verts;
  std::unordered_map<int, std::array<int, 2>> edges;
  std::vector<Hex> hexes;
  std::unordered_map<int, Curved> cedges, cfaces;
  std::map<std::pair<int, int>, int> edge_id;  // vertex pair -> edge id
  std::vector<double> coeffs;
  std::vector<int> coeff_ids;
  int ng = 0, nf = 0;
};

// The parse of the last size query on this thread.
thread_local std::unique_ptr<Parsed> t_parsed;

int parse_files(const char* xml_path, const char* fld_path, int flags, Parsed& P, std::string* err)
{
  auto bad = [&](int code, const std::string& m) {
    if (err) *err = "load_xml_fld: " + m;
    return code;
  };
  MappedFile xf, ff;
  if (!xf.open(xml_path)) return bad(FEMRT_EIO, std::string("cannot read ") + xml_path);
  if (fld_path && *fld_path && !ff.open(fld_path)) return bad(FEMRT_EIO, std::string("cannot read ") + fld_path);
  P.xml_path = xml_path;
  P.fld_path = fld_path ? fld_path : "";
  P.flags = flags;
  P.xml_size = (long long)xf.n;
  P.xml_mtime = xf.mtime_ns;
  P.fld_size = ff.fd >= 0 ? (long long)ff.n : -1;
  P.fld_mtime = ff.mtime_ns;
  const Text xml = xf.text();

  // ---- vertices, edges, faces
  std::string id;
  std::vector<double> v;
  each(section(xml, "VERTEX"), "V", [&](Text h, Text body) {
    if (!attr(h, "ID", id)) return;
    v.clear();
    parse_numbers(body, v);
    if (v.size() >= 3) P.verts[std::atoi(id.c_str())] = P3{v[0], v[1], v[2]};
  });
  if (P.verts.empty()) return bad(FEMRT_EIO, "no vertices");
  if ((flags & 1) && P.verts.count(0)) P.verts[0] = P3{};
  each(section(xml, "EDGE"), "E", [&](Text h, Text body) {
    if (!attr(h, "ID", id)) return;
    v.clear();
    parse_numbers(body, v);
    if (v.size() >= 2) P.edges[std::atoi(id.c_str())] = {(int)v[0], (int)v[1]};
  });
  std::unordered_map<int, std::array<int, 4>> face_verts;  // ring of vertices
  each(section(xml, "FACE"), "Q", [&](Text h, Text body) {
    if (!attr(h, "ID", id)) return;
    v.clear();
    parse_numbers(body, v);
    std::array<int, 4> ring;
    int nr = 0;
    for (double x : v) {
      const auto it = P.edges.find((int)x);
      if (it == P.edges.end()) return;
      for (int w : it->second) {
        bool seen = false;
        for (int k = 0; k < nr; ++k) seen = seen || ring[k] == w;
        if (!seen) {
          if (nr == 4) return;
          ring[nr++] = w;
        }
      }
    }
    if (nr == 4) face_verts[std::atoi(id.c_str())] = ring;
  });

  // ---- hexes
  const Text elem_sec = section(xml, "ELEMENT");
  for (const char* tag : {"H", "HEX"})
    each(elem_sec, tag, [&](Text h, Text body) {
      Hex x{};
      x.id = attr(h, "ID", id) ? std::atoi(id.c_str()) : (int)P.hexes.size();
      v.clear();
      parse_numbers(body, v);
      if (v.size() == 6) {
        x.has_faces = true;
        for (int k = 0; k < 6; ++k) x.f[k] = (int)v[k];
        for (int c = 0; c < 8; ++c) {
          x.v[c] = -1;
          const auto ia = face_verts.find(x.f[kCornerFaces[c][0]]);
          const auto ib = face_verts.find(x.f[kCornerFaces[c][1]]);
          const auto ic = face_verts.find(x.f[kCornerFaces[c][2]]);
          if (ia == face_verts.end() || ib == face_verts.end() || ic == face_verts.end()) continue;
          const auto &A = ia->second, &B = ib->second, &C = ic->second;
          for (int a : A)
            if (std::count(B.begin(), B.end(), a) && std::count(C.begin(), C.end(), a)) x.v[c] = a;
        }
      } else if (v.size() == 8) {
        x.has_faces = false;
        for (int c = 0; c < 8; ++c) x.v[c] = (int)v[c];
        x.f.fill(-1);
      } else {
        return;
      }
      P.hexes.push_back(x);
    });
  if (P.hexes.empty()) return bad(FEMRT_EIO, "no hexahedral elements");
  for (const Hex& x : P.hexes)
    for (int c = 0; c < 8; ++c)
      if (x.v[c] < 0 || !P.verts.count(x.v[c]))
        return bad(FEMRT_EIO, "element " + std::to_string(x.id) + ": unresolved corner vertex");
  std::sort(P.hexes.begin(), P.hexes.end(), [](const Hex& a, const Hex& b) { return a.id < b.id; });

  // ---- curved edges and faces
  const Text cur = section(xml, "CURVED");
  auto read_curved = [&](const char* tag, const char* key, std::unordered_map<int, Curved>& dst) {
    std::string np, ty;
    each(cur, tag, [&](Text h, Text body) {
      if (!attr(h, key, id) || !attr(h, "NUMPOINTS", np)) return;
      Curved c;
      c.npts = std::atoi(np.c_str());
      c.gll = attr(h, "TYPE", ty) && upper(ty).find("LOBATTO") != std::string::npos;
      v.clear();
      parse_numbers(body, v);
      c.pts.reserve(v.size() / 3);
      for (size_t k = 0; k + 2 < v.size(); k += 3) c.pts.push_back(P3{v[k], v[k + 1], v[k + 2]});
      if (!c.pts.empty()) dst[std::atoi(id.c_str())] = std::move(c);
    });
  };
  read_curved("E", "EDGEID", P.cedges);
  read_curved("F", "FACEID", P.cfaces);

  // ---- field modes: FLD NUMMODESPERDIR, else the XML expansion
  int fm[3] = {0, 0, 0};
  each(section(xml, "EXPANSIONS"), "E", [&](Text h, Text) {
    std::string m;
    if (!fm[0] && (attr(h, "NUMMODESPERDIR", m) || attr(h, "NUMMODES", m))) parse_modes(m, fm);
  });
  int fld_modes[3] = {0, 0, 0};
  if (ff.fd >= 0) {
    bool ok = true;
    each(ff.text(), "ELEMENTS", [&](Text h, Text body) {
      if (!ok) return;
      std::string m, ids, comp, bits, fields;
      int bm[3] = {0, 0, 0};
      if (attr(h, "NUMMODESPERDIR", m)) parse_modes(m, bm);
      if (!fld_modes[0]) std::copy(bm, bm + 3, fld_modes);
      const int nfield = attr(h, "FIELDS", fields) ? (int)std::count(fields.begin(), fields.end(), ',') + 1 : 1;
      std::vector<double> vals;
      // ASCII numbers, else base64 + zlib (files of older writers omit the
      // COMPRESSED attribute)
      bool ascii = !(attr(h, "COMPRESSED", comp) && upper(comp).find("B64Z") != std::string::npos);
      if (ascii) {
        const char* stop = nullptr;
        parse_numbers(body, vals, &stop);
        while (stop < body.e && is_space(*stop)) ++stop;
        if (stop < body.e) {
          ascii = false;
          vals.clear();
        }
      }
      if (!ascii) {
        std::vector<uint8_t> raw;
        if (!inflate_all(b64_decode(body), raw)) {
          ok = false;
          return;
        }
        const bool b32 = attr(h, "BITSIZE", bits) && std::atoi(bits.c_str()) == 32;
        const size_t w = b32 ? 4 : 8;
        vals.resize(raw.size() / w);
        for (size_t k = 0; k < vals.size(); ++k) {
          if (b32) {
            float f;
            std::memcpy(&f, &raw[k * 4], 4);
            vals[k] = f;
          } else {
            std::memcpy(&vals[k], &raw[k * 8], 8);
          }
        }
      }
      std::vector<int> idl = attr(h, "ID", ids) ? ranges(ids) : std::vector<int>();
      // first field of the block: element-major, n^3 coefficients each
      const size_t per = (size_t)(bm[0] ? bm[0] * bm[1] * bm[2] : 0);
      if (!idl.empty() && per && vals.size() >= per * idl.size() * nfield) {
        for (size_t k = 0; k < idl.size(); ++k) {
          P.coeff_ids.push_back(idl[k]);
          P.coeffs.insert(P.coeffs.end(), vals.begin() + k * per, vals.begin() + (k + 1) * per);
        }
      } else {
        P.coeffs.insert(P.coeffs.end(), vals.begin(), vals.end());
      }
    });
    if (!ok) return bad(FEMRT_EIO, "FLD payload is not valid base64 + zlib");
    if (fld_modes[0]) std::copy(fld_modes, fld_modes + 3, fm);
  }
  if (!fm[0]) return bad(FEMRT_EIO, "no field mode counts (EXPANSIONS or FLD)");
  if (fm[0] != fm[1] || fm[1] != fm[2]) return bad(FEMRT_EUNSUP, "anisotropic field orders");
  P.nf = fm[0];

  // ---- geometry modes
  for (const auto& kv : P.edges)
    P.edge_id[{std::min(kv.second[0], kv.second[1]), std::max(kv.second[0], kv.second[1])}] = kv.first;
  int ng = 2;
  for (const Hex& x : P.hexes) {
    if (x.has_faces)
      for (int f : x.f) {
        const auto it = P.cfaces.find(f);
        if (it != P.cfaces.end()) ng = std::max(ng, (int)std::lround(std::sqrt((double)it->second.npts)));
      }
    for (int e = 0; e < 12; ++e) {
      const int a = x.v[kEdges[e][0]], b = x.v[kEdges[e][1]];
      const auto ei = P.edge_id.find({std::min(a, b), std::max(a, b)});
      if (ei == P.edge_id.end()) continue;
      const auto ce = P.cedges.find(ei->second);
      if (ce != P.cedges.end()) ng = std::max(ng, (int)ce->second.pts.size());
    }
  }
  if (flags & 2) ng = std::max(ng, P.nf);
  if (ng > 16 || P.nf > 16) return bad(FEMRT_EUNSUP, "more than 16 modes per direction");
  P.ng = ng;
  return FEMRT_OK;
}

// Geometry map of element ie on the GLL grid, fitted to modes (the
// reference's BuildHexGeometry, xml_fld_geometry_builder.cpp:68-381).
void build_geometry(const Parsed& P, int ie, const std::vector<double>& gll, const std::vector<double>& inv,
                    std::vector<double>& nodal, std::vector<double>& coef, float* geom)
{
  const int ng = P.ng;
  const size_t g3 = (size_t)ng * ng * ng;
  const Hex& x = P.hexes[ie];
  P3 C[8];
  for (int c = 0; c < 8; ++c) C[c] = P.verts.at(x.v[c]);
  auto at = [&](int i, int j, int k) { return ((size_t)k * ng + j) * ng + i; };
  for (int k = 0; k < ng; ++k)
    for (int j = 0; j < ng; ++j)
      for (int i = 0; i < ng; ++i) {
        const double r = gll[i], s = gll[j], t = gll[k];
        double p[3] = {0, 0, 0};
        for (int c = 0; c < 8; ++c) {
          const double w = 0.125 * (1 + kCornerRst[c][0] * r) * (1 + kCornerRst[c][1] * s) *
                           (1 + kCornerRst[c][2] * t);
          p[0] += w * C[c].x;
          p[1] += w * C[c].y;
          p[2] += w * C[c].z;
        }
        for (int d = 0; d < 3; ++d) nodal[d * g3 + at(i, j, k)] = p[d];
      }
  // curved edges: interpolate the point set to the GLL nodes, in the
  // direction of increasing local parameter
  for (int e = 0; e < 12; ++e) {
    const int lo = x.v[kEdges[e][0]], hi = x.v[kEdges[e][1]];
    const auto ei = P.edge_id.find({std::min(lo, hi), std::max(lo, hi)});
    if (ei == P.edge_id.end()) continue;
    const auto ce = P.cedges.find(ei->second);
    if (ce == P.cedges.end() || (int)ce->second.pts.size() < 2) continue;
    const bool rev = P.edges.at(ei->second)[0] != lo;
    const int ns = (int)ce->second.pts.size();
    const std::vector<double> L = lagrange_matrix(nodes_of(ns, ce->second.gll), gll);
    const int ax = kEdges[e][2];
    const int* cl = kCornerRst[kEdges[e][0]];
    for (int m = 0; m < ng; ++m) {
      P3 q;
      for (int a = 0; a < ns; ++a) {
        const P3& s = ce->second.pts[rev ? ns - 1 - a : a];
        const double w = L[(size_t)m * ns + a];
        q.x += w * s.x;
        q.y += w * s.y;
        q.z += w * s.z;
      }
      int idx[3];
      for (int d = 0; d < 3; ++d) idx[d] = cl[d] < 0 ? 0 : ng - 1;
      idx[ax] = m;
      const size_t id = at(idx[0], idx[1], idx[2]);
      nodal[id] = q.x;
      nodal[g3 + id] = q.y;
      nodal[2 * g3 + id] = q.z;
    }
  }
  // curved faces: orient the point grid to the local face, interpolate to
  // the GLL grid, blend the difference along the normal coordinate
  if (x.has_faces)
    for (int lf = 0; lf < 6; ++lf) {
      const auto cf = P.cfaces.find(x.f[lf]);
      if (cf == P.cfaces.end()) continue;
      const int ns = (int)std::lround(std::sqrt((double)cf->second.pts.size()));
      if (ns * ns != (int)cf->second.pts.size() || ns < 2) continue;
      const auto& G = cf->second.pts;
      P3 want[4];
      for (int c = 0; c < 4; ++c) want[c] = C[kFaceCorners[lf][c]];
      // the one of the 8 grid symmetries whose corners match the face's
      int best = -1;
      double best_err = 1e300;
      for (int sym = 0; sym < 8; ++sym) {
        auto src = [&](int i, int j) {
          int a = (sym & 1) ? j : i, b = (sym & 1) ? i : j;
          if (sym & 2) a = ns - 1 - a;
          if (sym & 4) b = ns - 1 - b;
          return G[(size_t)b * ns + a];
        };
        const double e = dist2(src(0, 0), want[0]) + dist2(src(ns - 1, 0), want[1]) +
                         dist2(src(ns - 1, ns - 1), want[2]) + dist2(src(0, ns - 1), want[3]);
        if (e < best_err) {
          best_err = e;
          best = sym;
        }
      }
      const std::vector<double> L = lagrange_matrix(nodes_of(ns, cf->second.gll), gll);
      const int fix = (lf == 0 || lf == 5) ? 2 : ((lf == 1 || lf == 3) ? 1 : 0);
      const bool plus = lf == 2 || lf == 3 || lf == 5;
      const int ua = (lf == 2 || lf == 4) ? 1 : 0, va = (lf == 0 || lf == 5) ? 1 : 2;
      for (int b2 = 0; b2 < ng; ++b2)
        for (int a2 = 0; a2 < ng; ++a2) {
          P3 q;
          for (int b = 0; b < ns; ++b)
            for (int a = 0; a < ns; ++a) {
              int sa = (best & 1) ? b : a, sb = (best & 1) ? a : b;
              if (best & 2) sa = ns - 1 - sa;
              if (best & 4) sb = ns - 1 - sb;
              const P3& s = G[(size_t)sb * ns + sa];
              const double w = L[(size_t)a2 * ns + a] * L[(size_t)b2 * ns + b];
              q.x += w * s.x;
              q.y += w * s.y;
              q.z += w * s.z;
            }
          int idx[3];
          idx[ua] = a2;
          idx[va] = b2;
          idx[fix] = plus ? ng - 1 : 0;
          const size_t fid = at(idx[0], idx[1], idx[2]);
          const double diff[3] = {q.x - nodal[fid], q.y - nodal[g3 + fid], q.z - nodal[2 * g3 + fid]};
          for (int m = 0; m < ng; ++m) {
            idx[fix] = m;
            const double w = plus ? 0.5 * (1 + gll[m]) : 0.5 * (1 - gll[m]);
            const size_t id = at(idx[0], idx[1], idx[2]);
            for (int d = 0; d < 3; ++d) nodal[d * g3 + id] += diff[d] * w;
          }
        }
    }
  for (int d = 0; d < 3; ++d) {
    tensor_fit(ng, inv, &nodal[d * g3], coef.data());
    for (size_t m = 0; m < g3; ++m) geom[((size_t)ie * 3 + d) * g3 + m] = (float)coef[m];
  }
}

// Runs f(lo, hi) over [0, n) split across the host's cores.
template <typename F>
void parallel_ranges(int n, F f)
{
  const int hw = (int)std::max(1u, std::thread::hardware_concurrency());
  const int T = std::max(1, std::min({hw, 16, n / 256 + 1}));
  if (T == 1) {
    f(0, n);
    return;
  }
  std::vector<std::thread> th;
  for (int t = 0; t < T; ++t) th.emplace_back(f, (int)((long long)n * t / T), (int)((long long)n * (t + 1) / T));
  for (auto& x : th) x.join();
}

}  // namespace

// Two-call protocol: with every output pointer NULL only the sizes are
// returned (the parse is kept for the next call on this thread with the same
// files, flags and file times); otherwise the outputs are filled.
int load_xml_fld(const char* xml_path, const char* fld_path, int flags, int* n_elem_out, int* ng_out,
                 int* nf_out, float* vertices, float* geom, float* field, int* adjacency,
                 std::string* err)
{
  if (!xml_path || !n_elem_out || !ng_out || !nf_out) {
    if (err) *err = "load_xml_fld: null argument";
    return FEMRT_EINVAL;
  }
  const bool sizes_only = !vertices && !geom && !field && !adjacency;
  std::unique_ptr<Parsed> P = std::move(t_parsed);
  bool reuse = false;
  if (P) {
    struct stat sx, sf;
    const std::string fp = fld_path ? fld_path : "";
    reuse = P->xml_path == xml_path && P->fld_path == fp && P->flags == flags && stat(xml_path, &sx) == 0 &&
            (long long)sx.st_size == P->xml_size &&
            (long long)sx.st_mtim.tv_sec * 1000000000LL + sx.st_mtim.tv_nsec == P->xml_mtime &&
            (fp.empty() ? P->fld_size < 0
                        : (stat(fp.c_str(), &sf) == 0 && (long long)sf.st_size == P->fld_size &&
                           (long long)sf.st_mtim.tv_sec * 1000000000LL + sf.st_mtim.tv_nsec == P->fld_mtime));
  }
  if (!reuse) {
    P.reset(new Parsed);
    const int rc = parse_files(xml_path, fld_path, flags, *P, err);
    if (rc != FEMRT_OK) return rc;
  }
  const int n = (int)P->hexes.size(), ng = P->ng, nf = P->nf;
  *n_elem_out = n;
  *ng_out = ng;
  *nf_out = nf;
  if (sizes_only) {
    t_parsed = std::move(P);
    return FEMRT_OK;
  }

  // ---- per-element output, on all host cores
  std::unordered_map<int, size_t> coeff_pos;
  for (size_t k = 0; k < P->coeff_ids.size(); ++k) coeff_pos[P->coeff_ids[k]] = k;
  const size_t f3 = (size_t)nf * nf * nf, g3 = (size_t)ng * ng * ng;
  const std::vector<double> gll = gll_nodes(ng);
  std::vector<double> inv;
  modal_fit_1d_matrix(ng, gll, inv);
  const Parsed& Q = *P;
  parallel_ranges(n, [&](int lo, int hi) {
    std::vector<double> nodal(3 * g3), coef(g3);
    for (int ie = lo; ie < hi; ++ie) {
      const Hex& x = Q.hexes[ie];
      if (vertices)
        for (int c = 0; c < 8; ++c) {
          const P3& C = Q.verts.at(x.v[c]);
          vertices[(size_t)ie * 24 + 3 * c + 0] = (float)C.x;
          vertices[(size_t)ie * 24 + 3 * c + 1] = (float)C.y;
          vertices[(size_t)ie * 24 + 3 * c + 2] = (float)C.z;
        }
      if (field) {
        size_t pos = (size_t)-1;
        if (!Q.coeff_ids.empty()) {
          const auto it = coeff_pos.find(x.id);
          if (it != coeff_pos.end()) pos = it->second * f3;
        } else if ((size_t)(ie + 1) * f3 <= Q.coeffs.size()) {
          pos = (size_t)ie * f3;  // expansion order
        }
        for (size_t m = 0; m < f3; ++m)
          field[(size_t)ie * f3 + m] = pos == (size_t)-1 ? 0.f : (float)Q.coeffs[pos + m];
      }
      if (geom) build_geometry(Q, ie, gll, inv, nodal, coef, geom);
    }
  });

  // ---- adjacency: elements sharing a face (by its 4 vertices)
  if (adjacency) {
    std::vector<std::pair<std::array<int, 4>, int>> keys((size_t)n * 6);  // face vertices, ie * 6 + lf
    for (int ie = 0; ie < n; ++ie)
      for (int lf = 0; lf < 6; ++lf) {
        std::array<int, 4> key;
        for (int c = 0; c < 4; ++c) key[c] = Q.hexes[ie].v[kFaceCorners[lf][c]];
        std::sort(key.begin(), key.end());
        keys[(size_t)ie * 6 + lf] = {key, ie * 6 + lf};
        adjacency[((size_t)ie * 6 + lf) * 2] = -1;
        adjacency[((size_t)ie * 6 + lf) * 2 + 1] = -1;
      }
    std::sort(keys.begin(), keys.end());
    for (size_t k = 0; k < keys.size();) {
      size_t j = k + 1;
      while (j < keys.size() && keys[j].first == keys[k].first) ++j;
      if (j - k == 2) {
        const int a = keys[k].second, b = keys[k + 1].second;
        adjacency[(size_t)a * 2] = b / 6;
        adjacency[(size_t)a * 2 + 1] = b % 6;
        adjacency[(size_t)b * 2] = a / 6;
        adjacency[(size_t)b * 2 + 1] = a % 6;
      }
      k = j;
    }
  }
  return FEMRT_OK;
}

// Writer of a structured or general uniform-order hex mesh: vertices and
// edges deduplicated by position, every face written as CURVED with its
// (ng)^2 evenly spaced points from the modal geometry map; FLD as one
// ELEMENTS block, B64Z-LittleEndian 64-bit, NUMMODESPERDIR="UNIORDER:n,n,n".
int save_xml_fld(const char* xml_path, const char* fld_path, int n, int ng, int nf,
                 const float* vertices, const float* geom, const float* field, std::string* err)
{
  auto bad = [&](int code, const std::string& m) {
    if (err) *err = "save_xml_fld: " + m;
    return code;
  };
  if (!xml_path || !fld_path || n <= 0 || ng < 2 || nf < 1 || !vertices || !geom || !field)
    return bad(FEMRT_EINVAL, "bad arguments");
  // vertex ids by exact float position
  std::map<std::array<float, 3>, int> vid;
  std::vector<std::array<float, 3>> vpos;
  std::vector<std::array<int, 8>> ev(n);
  for (int e = 0; e < n; ++e)
    for (int c = 0; c < 8; ++c) {
      const std::array<float, 3> p = {vertices[(size_t)e * 24 + 3 * c], vertices[(size_t)e * 24 + 3 * c + 1],
                                      vertices[(size_t)e * 24 + 3 * c + 2]};
      auto it = vid.find(p);
      if (it == vid.end()) {
        it = vid.emplace(p, (int)vpos.size()).first;
        vpos.push_back(p);
      }
      ev[e][c] = it->second;
    }
  std::map<std::pair<int, int>, int> eid;
  std::vector<std::pair<int, int>> elist;
  std::map<std::array<int, 4>, int> fid;
  std::vector<std::array<int, 4>> flist_edges;
  std::vector<std::pair<int, int>> fowner;  // element, local face
  std::vector<std::array<int, 6>> hf(n);
  const int fe[6][4][2] = {{{0, 1}, {1, 2}, {2, 3}, {3, 0}}, {{0, 1}, {1, 5}, {5, 4}, {4, 0}},
                           {{1, 2}, {2, 6}, {6, 5}, {5, 1}}, {{3, 2}, {2, 6}, {6, 7}, {7, 3}},
                           {{0, 3}, {3, 7}, {7, 4}, {4, 0}}, {{4, 5}, {5, 6}, {6, 7}, {7, 4}}};
  for (int e = 0; e < n; ++e)
    for (int lf = 0; lf < 6; ++lf) {
      std::array<int, 4> ids;
      for (int k = 0; k < 4; ++k) {
        const int a = ev[e][fe[lf][k][0]], b = ev[e][fe[lf][k][1]];
        const std::pair<int, int> key{std::min(a, b), std::max(a, b)};
        auto it = eid.find(key);
        if (it == eid.end()) {
          it = eid.emplace(key, (int)elist.size()).first;
          elist.push_back({a, b});
        }
        ids[k] = it->second;
      }
      std::array<int, 4> key = ids;
      std::sort(key.begin(), key.end());
      auto it = fid.find(key);
      if (it == fid.end()) {
        it = fid.emplace(key, (int)flist_edges.size()).first;
        flist_edges.push_back(ids);
        fowner.push_back({e, lf});
      }
      hf[e][lf] = it->second;
    }
  FILE* fx = std::fopen(xml_path, "w");
  if (!fx) return bad(FEMRT_EIO, std::string("cannot write ") + xml_path);
  std::fprintf(fx, "<?xml version=\"1.0\" encoding=\"utf-8\" ?>\n<NEKTAR>\n  <GEOMETRY DIM=\"3\" SPACE=\"3\">\n    <VERTEX>\n");
  for (size_t v = 0; v < vpos.size(); ++v)
    std::fprintf(fx, "      <V ID=\"%zu\"> %.9g %.9g %.9g </V>\n", v, vpos[v][0], vpos[v][1], vpos[v][2]);
  std::fprintf(fx, "    </VERTEX>\n    <EDGE>\n");
  for (size_t k = 0; k < elist.size(); ++k)
    std::fprintf(fx, "      <E ID=\"%zu\"> %d %d </E>\n", k, elist[k].first, elist[k].second);
  std::fprintf(fx, "    </EDGE>\n    <FACE>\n");
  for (size_t k = 0; k < flist_edges.size(); ++k)
    std::fprintf(fx, "      <Q ID=\"%zu\"> %d %d %d %d </Q>\n", k, flist_edges[k][0], flist_edges[k][1],
                 flist_edges[k][2], flist_edges[k][3]);
  std::fprintf(fx, "    </FACE>\n    <ELEMENT>\n");
  for (int e = 0; e < n; ++e)
    std::fprintf(fx, "      <H ID=\"%d\"> %d %d %d %d %d %d </H>\n", e, hf[e][0], hf[e][1], hf[e][2],
                 hf[e][3], hf[e][4], hf[e][5]);
  std::fprintf(fx, "    </ELEMENT>\n    <CURVED>\n");
  // face points from the owning element's modal map, on the face's own
  // (u, v) grid: point (a, b) at u = -1 + 2a/(ng-1), v = -1 + 2b/(ng-1);
  // the faces are evaluated on all host cores into text blocks written in
  // face order
  const size_t g3 = (size_t)ng * ng * ng;
  std::vector<double> phi((size_t)(ng + 2) * ng);  // phi[x][i]: x = grid point a, ng: -1, ng + 1: +1
  for (int a = 0; a < ng + 2; ++a) {
    const double xa = a < ng ? -1.0 + 2.0 * a / (ng - 1) : (a == ng ? -1.0 : 1.0);
    for (int i = 0; i < ng; ++i) phi[(size_t)a * ng + i] = modA(i, xa);
  }
  const int nfaces = (int)fowner.size();
  std::vector<std::string> ftext(nfaces);
  parallel_ranges(nfaces, [&](int lo, int hi) {
    char buf[96];
    for (int k = lo; k < hi; ++k) {
      const int e = fowner[k].first, lf = fowner[k].second;
      std::string& out = ftext[k];
      std::snprintf(buf, sizeof buf, "      <F ID=\"%d\" FACEID=\"%d\" NUMPOINTS=\"%d\" TYPE=\"PolyEvenlySpaced\">",
                    k, k, ng * ng);
      out += buf;
      const int fix = (lf == 0 || lf == 5) ? 2 : ((lf == 1 || lf == 3) ? 1 : 0);
      const int fv = (lf == 2 || lf == 3 || lf == 5) ? ng + 1 : ng;
      const int ua = (lf == 2 || lf == 4) ? 1 : 0, va = (lf == 0 || lf == 5) ? 1 : 2;
      for (int b = 0; b < ng; ++b)
        for (int a = 0; a < ng; ++a) {
          int row[3];
          row[fix] = fv;
          row[ua] = a;
          row[va] = b;
          const double* pr = &phi[(size_t)row[0] * ng];
          const double* ps = &phi[(size_t)row[1] * ng];
          const double* pt = &phi[(size_t)row[2] * ng];
          double x[3] = {0, 0, 0};
          for (int kk = 0; kk < ng; ++kk)
            for (int jj = 0; jj < ng; ++jj) {
              const double wjk = ps[jj] * pt[kk];
              for (int ii = 0; ii < ng; ++ii) {
                const double w = pr[ii] * wjk;
                const size_t m = ((size_t)kk * ng + jj) * ng + ii;
                for (int d = 0; d < 3; ++d) x[d] += w * geom[((size_t)e * 3 + d) * g3 + m];
              }
            }
          std::snprintf(buf, sizeof buf, " %.9g %.9g %.9g", x[0], x[1], x[2]);
          out += buf;
        }
      out += " </F>\n";
    }
  });
  for (const std::string& t : ftext) std::fwrite(t.data(), 1, t.size(), fx);
  std::fprintf(fx, "    </CURVED>\n    <COMPOSITE>\n      <C ID=\"0\"> H[0-%d] </C>\n    </COMPOSITE>\n", n - 1);
  std::fprintf(fx, "    <DOMAIN> C[0] </DOMAIN>\n  </GEOMETRY>\n  <EXPANSIONS>\n");
  std::fprintf(fx, "    <E COMPOSITE=\"C[0]\" NUMMODES=\"%d\" FIELDS=\"u\" TYPE=\"MODIFIED\" />\n", nf);
  std::fprintf(fx, "  </EXPANSIONS>\n</NEKTAR>\n");
  const bool xml_ok = std::fclose(fx) == 0;
  if (!xml_ok) return bad(FEMRT_EIO, std::string("cannot write ") + xml_path);
  // FLD
  const size_t f3 = (size_t)nf * nf * nf;
  std::vector<double> d((size_t)n * f3);
  for (size_t k = 0; k < d.size(); ++k) d[k] = field[k];
  uLongf zlen = compressBound((uLong)(d.size() * 8));
  std::vector<uint8_t> z(zlen);
  if (compress2(z.data(), &zlen, (const Bytef*)d.data(), (uLong)(d.size() * 8), Z_BEST_SPEED) != Z_OK)
    return bad(FEMRT_EIO, "zlib compression failed");
  FILE* ff = std::fopen(fld_path, "w");
  if (!ff) return bad(FEMRT_EIO, std::string("cannot write ") + fld_path);
  std::fprintf(ff, "<?xml version=\"1.0\" encoding=\"utf-8\" ?>\n<NEKTAR>\n");
  std::fprintf(ff, "    <ELEMENTS FIELDS=\"u\" SHAPE=\"Hexahedron\" BASIS=\"Modified_A,Modified_A,Modified_A\" "
               "NUMMODESPERDIR=\"UNIORDER:%d,%d,%d\" ID=\"0-%d\" COMPRESSED=\"B64Z-LittleEndian\" BITSIZE=\"64\">",
               nf, nf, nf, n - 1);
  const std::string enc = b64_encode(z.data(), zlen);
  std::fwrite(enc.data(), 1, enc.size(), ff);
  std::fprintf(ff, "</ELEMENTS>\n</NEKTAR>\n");
  if (std::fclose(ff) != 0) return bad(FEMRT_EIO, std::string("cannot write ") + fld_path);
  return FEMRT_OK;
}

}  // namespace femrt
