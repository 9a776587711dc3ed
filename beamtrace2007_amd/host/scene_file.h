// Resident scene format (.btgs): a scene and its kd-tree as the GPU path
// reads them, so that a server (or a second run) uploads a scene without
// regenerating it or rebuilding the tree. The reference has no such format:
// it rebuilds the kd-tree on every start (KdTree::build,
// src/accel/kdtree.cpp:237-281), which for a 2M-triangle scene takes seconds.
//
// Layout, little-endian, every section starting at a multiple of 8 bytes:
//   header (256 bytes)
//     0  char[8]  magic "BTGSCENE"
//     8  u32      version (1)
//    12  u32      spec flags (1 point light, 2 area light, 4 cull back faces)
//    16  i64 x 4  numVerts, numTris, numNodes (0: no kd-tree), numRefs
//    48  i32      kd depth
//    52  u32      reserved (0)
//    56  f32 x 6  scene bounds (lo, hi)
//    80  f32 x 6  kd root box (lo, hi)
//   104  f32 x 10 view (eye, target, up, vertical fov in degrees)
//   144  f32 x 3  point light
//   156  f32 x 9  area light (centre, U, V)
//   192  u64      payload bytes
//   200  u64      payload hash (64-bit FNV-1a over the payload's 8-byte words)
//   208  zero to 256
//   payload sections, each zero-padded to 8 bytes:
//     pos f32[3 numVerts], tri3 u32[3 numTris], triN f32[3 numTris],
//     triArea f32[numTris], nodes 12 B[numNodes] {f32 split, u32 data,
//     u32 offset}, triIndices u32[numRefs], parent u32[numNodes],
//     nodeBox f32[6 numNodes]
#ifndef BTH_SCENE_FILE_H
#define BTH_SCENE_FILE_H

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace btsf {

constexpr char kMagic[8] = {'B', 'T', 'G', 'S', 'C', 'E', 'N', 'E'};
constexpr uint32_t kVersion = 1;
constexpr size_t kHeaderBytes = 256;

struct Header {
  uint32_t specFlags = 0;
  int64_t numVerts = 0, numTris = 0, numNodes = 0, numRefs = 0;
  int32_t kdDepth = 0;
  float sceneBounds[6] = {}, kdBounds[6] = {};
  float view[10] = {}, point[3] = {}, light[9] = {};
};

// Borrowed arrays of a scene to write.
struct View {
  Header h;
  const float* pos = nullptr;
  const uint32_t* tri3 = nullptr;
  const float* triN = nullptr;
  const float* triArea = nullptr;
  const void* nodes = nullptr;
  const uint32_t* triIndices = nullptr;
  const uint32_t* parent = nullptr;
  const float* nodeBox = nullptr;
};

// Owned arrays of a scene read back.
struct Data {
  Header h;
  std::vector<float> pos, triN, triArea, nodeBox;
  std::vector<uint32_t> tri3, triIndices, parent;
  std::vector<uint8_t> nodes;
};

inline size_t pad8(size_t n) { return (n + 7) & ~size_t(7); }

inline uint64_t hashWords(uint64_t h, const uint8_t* p, size_t n) {  // n: a multiple of 8
  for (size_t i = 0; i < n; i += 8) {
    uint64_t w;
    std::memcpy(&w, p + i, 8);
    h ^= w;
    h *= 1099511628211ull;
  }
  return h;
}

inline void sectionSizes(const Header& h, size_t (&s)[8]) {
  s[0] = size_t(h.numVerts) * 12;
  s[1] = size_t(h.numTris) * 12;
  s[2] = size_t(h.numTris) * 12;
  s[3] = size_t(h.numTris) * 4;
  s[4] = size_t(h.numNodes) * 12;
  s[5] = size_t(h.numRefs) * 4;
  s[6] = size_t(h.numNodes) * 4;
  s[7] = size_t(h.numNodes) * 24;
}

inline void putHeader(const Header& h, uint64_t payloadBytes, uint64_t hash, uint8_t (&b)[kHeaderBytes]) {
  std::memset(b, 0, sizeof b);
  std::memcpy(b, kMagic, 8);
  std::memcpy(b + 8, &kVersion, 4);
  std::memcpy(b + 12, &h.specFlags, 4);
  const int64_t counts[4] = {h.numVerts, h.numTris, h.numNodes, h.numRefs};
  std::memcpy(b + 16, counts, 32);
  std::memcpy(b + 48, &h.kdDepth, 4);
  std::memcpy(b + 56, h.sceneBounds, 24);
  std::memcpy(b + 80, h.kdBounds, 24);
  std::memcpy(b + 104, h.view, 40);
  std::memcpy(b + 144, h.point, 12);
  std::memcpy(b + 156, h.light, 36);
  std::memcpy(b + 192, &payloadBytes, 8);
  std::memcpy(b + 200, &hash, 8);
}

inline bool write(const char* path, const View& v, std::string& err) {
  size_t s[8];
  sectionSizes(v.h, s);
  const void* src[8] = {v.pos, v.tri3, v.triN, v.triArea, v.nodes, v.triIndices, v.parent, v.nodeBox};
  uint64_t hash = 1469598103934665603ull, total = 0;
  const uint8_t zero[8] = {};
  for (int k = 0; k < 8; ++k) {
    if (s[k] && !src[k]) {
      err = "scene file: missing array";
      return false;
    }
    const size_t whole = s[k] & ~size_t(7);
    if (whole) hash = hashWords(hash, static_cast<const uint8_t*>(src[k]), whole);
    if (s[k] != whole) {
      uint8_t tail[8] = {};
      std::memcpy(tail, static_cast<const uint8_t*>(src[k]) + whole, s[k] - whole);
      hash = hashWords(hash, tail, 8);
    }
    total += pad8(s[k]);
  }
  uint8_t hb[kHeaderBytes];
  putHeader(v.h, total, hash, hb);
  const std::string tmp = std::string(path) + ".tmp";
  FILE* f = std::fopen(tmp.c_str(), "wb");
  if (!f) {
    err = "scene file: cannot open " + tmp + " for writing";
    return false;
  }
  bool ok = std::fwrite(hb, 1, sizeof hb, f) == sizeof hb;
  for (int k = 0; k < 8 && ok; ++k) {
    if (s[k]) ok = std::fwrite(src[k], 1, s[k], f) == s[k];
    if (ok && pad8(s[k]) != s[k]) ok = std::fwrite(zero, 1, pad8(s[k]) - s[k], f) == pad8(s[k]) - s[k];
  }
  ok = (std::fclose(f) == 0) && ok;
  if (!ok || std::rename(tmp.c_str(), path) != 0) {
    std::remove(tmp.c_str());
    err = std::string("scene file: write to ") + path + " failed";
    return false;
  }
  return true;
}

inline bool read(const char* path, Data& d, std::string& err) {
  FILE* f = std::fopen(path, "rb");
  if (!f) {
    err = std::string("scene file: cannot open ") + path;
    return false;
  }
  auto fail = [&](const std::string& m) {
    std::fclose(f);
    err = std::string("scene file ") + path + ": " + m;
    return false;
  };
  uint8_t b[kHeaderBytes];
  if (std::fread(b, 1, sizeof b, f) != sizeof b) return fail("truncated header");
  if (std::memcmp(b, kMagic, 8) != 0) return fail("not a scene file");
  uint32_t version;
  std::memcpy(&version, b + 8, 4);
  if (version != kVersion) return fail("unsupported version " + std::to_string(version));
  Header& h = d.h;
  std::memcpy(&h.specFlags, b + 12, 4);
  int64_t counts[4];
  std::memcpy(counts, b + 16, 32);
  h.numVerts = counts[0];
  h.numTris = counts[1];
  h.numNodes = counts[2];
  h.numRefs = counts[3];
  std::memcpy(&h.kdDepth, b + 48, 4);
  std::memcpy(h.sceneBounds, b + 56, 24);
  std::memcpy(h.kdBounds, b + 80, 24);
  std::memcpy(h.view, b + 104, 40);
  std::memcpy(h.point, b + 144, 12);
  std::memcpy(h.light, b + 156, 36);
  uint64_t payload, hash;
  std::memcpy(&payload, b + 192, 8);
  std::memcpy(&hash, b + 200, 8);
  for (int64_t c : counts)
    if (c < 0 || c >= (int64_t(1) << 32)) return fail("count out of range");
  size_t s[8];
  sectionSizes(h, s);
  uint64_t total = 0;
  for (size_t x : s) total += pad8(x);
  if (total != payload) return fail("payload size does not match the counts");
  d.pos.resize(size_t(h.numVerts) * 3);
  d.tri3.resize(size_t(h.numTris) * 3);
  d.triN.resize(size_t(h.numTris) * 3);
  d.triArea.resize(size_t(h.numTris));
  d.nodes.resize(size_t(h.numNodes) * 12);
  d.triIndices.resize(size_t(h.numRefs));
  d.parent.resize(size_t(h.numNodes));
  d.nodeBox.resize(size_t(h.numNodes) * 6);
  void* dst[8] = {d.pos.data(), d.tri3.data(), d.triN.data(), d.triArea.data(),
                  d.nodes.data(), d.triIndices.data(), d.parent.data(), d.nodeBox.data()};
  uint64_t got = 1469598103934665603ull;
  for (int k = 0; k < 8; ++k) {
    if (s[k] && std::fread(dst[k], 1, s[k], f) != s[k]) return fail("truncated payload");
    uint8_t padb[8] = {};
    const size_t extra = pad8(s[k]) - s[k];
    if (extra && std::fread(padb, 1, extra, f) != extra) return fail("truncated payload");
    const size_t whole = s[k] & ~size_t(7);
    if (whole) got = hashWords(got, static_cast<const uint8_t*>(dst[k]), whole);
    if (extra) {
      uint8_t tail[8] = {};
      std::memcpy(tail, static_cast<const uint8_t*>(dst[k]) + whole, s[k] - whole);
      got = hashWords(got, tail, 8);
    }
  }
  if (std::fgetc(f) != EOF) return fail("trailing bytes after the payload");
  std::fclose(f);
  if (got != hash) {
    err = std::string("scene file ") + path + ": payload hash mismatch (corrupt file)";
    return false;
  }
  return true;
}

}  // namespace btsf

#endif
