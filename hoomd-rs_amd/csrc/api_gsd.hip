// GSD trajectory files through the C ABI: the GSD 2.1 file layer
// (hoomd-gsd/src/file_layer.rs) and HOOMD schema frames of a context's state
// (hoomd-gsd/src/hoomd.rs).
//
// Layout (native byte order): a 256-byte header (magic, index location and
// capacity in entries, name-list location and capacity in 64-byte names,
// schema version (major << 16 | minor), file version (2 << 16 | 1), the
// application name at byte 48 and the schema name at byte 112, each at most
// 63 bytes and zero-terminated); the index, 32-byte entries (frame, rows,
// data location, columns, name id, type, flags) sorted by (frame, id), unused
// entries zero; the name list, zero-terminated names in 64-byte slots in id
// order; and the chunk data.  A new file holds the header, a 128-entry index
// and a 16-name list.  A full index or name list moves to the end of the file
// with twice the capacity.
#include "host.hpp"

#include <algorithm>
#include <cstdio>
#include <set>
#include <string>
#include <unordered_map>

namespace {

constexpr uint64_t kGsdMagic = 0x65df65df65df65dfull;
constexpr uint64_t kHeaderBytes = 256, kEntryBytes = 32, kNameBytes = 64;
constexpr uint64_t kFirstEntries = 128, kFirstNames = 1024 / kNameBytes;
constexpr uint32_t kFileVersion = (2u << 16) | 1u;

struct GsdHeader {
  uint64_t magic, index_location, index_entries, names_location, names_entries;
  uint32_t schema_version, file_version;
  char application[64];
  char schema[64];
  char reserved[80];
};
static_assert(sizeof(GsdHeader) == kHeaderBytes, "GSD header size");

struct GsdEntry {
  uint64_t frame, rows, location;
  uint32_t columns;
  uint16_t id;
  uint8_t type, flags;
};
static_assert(sizeof(GsdEntry) == kEntryBytes, "GSD index entry size");

uint64_t type_bytes(uint8_t t) {
  switch (t) {
    case HMC_GSD_U8: case HMC_GSD_I8: case HMC_GSD_STRING: return 1;
    case HMC_GSD_U16: case HMC_GSD_I16: return 2;
    case HMC_GSD_U32: case HMC_GSD_I32: case HMC_GSD_F32: return 4;
    case HMC_GSD_U64: case HMC_GSD_I64: case HMC_GSD_F64: return 8;
    default: return 0;
  }
}

bool entry_less(const GsdEntry& a, const GsdEntry& b) { return a.frame != b.frame ? a.frame < b.frame : a.id < b.id; }

}  // namespace

struct hmc_gsd {
  FILE* file = nullptr;
  bool writable = false;
  GsdHeader h{};
  std::vector<std::string> names;                  // by id
  std::unordered_map<std::string, uint16_t> ids;
  std::vector<GsdEntry> index;                     // entries in the file, sorted
  uint64_t frames = 0;                             // completed frames
  uint64_t file_len = 0;
  // the frame being written
  std::vector<GsdEntry> pending;
  std::vector<uint8_t> data;
  size_t names_written = 0;                        // names already in the file's list
  std::set<uint16_t> frame_ids;
};

namespace {

bool put(hmc_gsd* f, uint64_t at, const void* p, size_t n) {
  return std::fseek(f->file, (long)at, SEEK_SET) == 0 && std::fwrite(p, 1, n, f->file) == n;
}
bool get(const hmc_gsd* f, uint64_t at, void* p, size_t n) {
  return std::fseek(f->file, (long)at, SEEK_SET) == 0 && std::fread(p, 1, n, f->file) == n;
}
// a zero-filled block of n bytes at the end of the file; returns its location
bool grow(hmc_gsd* f, uint64_t n, uint64_t* at) {
  *at = f->file_len;
  std::vector<uint8_t> zero(std::min<uint64_t>(n, 1u << 20), 0);
  for (uint64_t done = 0; done < n;) {
    const size_t k = (size_t)std::min<uint64_t>(zero.size(), n - done);
    if (!put(f, f->file_len + done, zero.data(), k)) return false;
    done += k;
  }
  f->file_len += n;
  return true;
}
// the smallest doubling of `have` entries above `need`
uint64_t doubled(uint64_t have, uint64_t need) {
  uint64_t c = std::max<uint64_t>(have, 1);
  while (c <= need) c *= 2;
  return c;
}

// Writes the frame in progress: data, then new names and index entries
// (moving the name list or the index to the end of the file when full),
// then the header when a block moved.
bool flush_frame(hmc_gsd* f) {
  bool header_dirty = false;
  if (!f->data.empty()) {
    if (!put(f, f->file_len, f->data.data(), f->data.size())) return false;
    f->file_len += f->data.size();
    f->data.clear();
  }
  if (f->names.size() > f->names_written) {
    if (f->names.size() > f->h.names_entries) {
      const uint64_t cap = doubled(f->h.names_entries, f->names.size());
      uint64_t at = 0;
      if (!grow(f, cap * kNameBytes, &at)) return false;
      f->h.names_location = at;
      f->h.names_entries = cap;
      f->names_written = 0;  // rewrite every name in the new list
      header_dirty = true;
    }
    for (size_t k = f->names_written; k < f->names.size(); ++k) {
      char slot[kNameBytes] = {};
      std::memcpy(slot, f->names[k].data(), f->names[k].size());
      if (!put(f, f->h.names_location + k * kNameBytes, slot, kNameBytes)) return false;
    }
    f->names_written = f->names.size();
  }
  if (!f->pending.empty()) {
    std::sort(f->pending.begin(), f->pending.end(), entry_less);
    const uint64_t need = f->index.size() + f->pending.size();
    uint64_t from = f->index.size();
    if (need > f->h.index_entries) {
      const uint64_t cap = doubled(f->h.index_entries, need);
      uint64_t at = 0;
      if (!grow(f, cap * kEntryBytes, &at)) return false;
      f->h.index_location = at;
      f->h.index_entries = cap;
      from = 0;  // rewrite every entry in the new index
      header_dirty = true;
    }
    f->index.insert(f->index.end(), f->pending.begin(), f->pending.end());
    f->pending.clear();
    if (!put(f, f->h.index_location + from * kEntryBytes, f->index.data() + from,
             (f->index.size() - from) * kEntryBytes))
      return false;
  }
  if (header_dirty && !put(f, 0, &f->h, kHeaderBytes)) return false;
  return std::fflush(f->file) == 0;
}

bool load(hmc_gsd* f) {
  if (!get(f, 0, &f->h, kHeaderBytes) || f->h.magic != kGsdMagic) return false;
  if ((f->h.file_version >> 16) != 2u) return false;  // 2.x files only
  f->h.application[63] = 0;
  f->h.schema[63] = 0;
  if (std::fseek(f->file, 0, SEEK_END) != 0) return false;
  f->file_len = (uint64_t)std::ftell(f->file);
  if (f->h.names_location + f->h.names_entries * kNameBytes > f->file_len ||
      f->h.index_location + f->h.index_entries * kEntryBytes > f->file_len)
    return false;
  std::vector<char> nl(f->h.names_entries * kNameBytes);
  if (!nl.empty() && !get(f, f->h.names_location, nl.data(), nl.size())) return false;
  for (uint64_t k = 0; k < f->h.names_entries && nl[k * kNameBytes] != 0; ++k) {
    const char* s = &nl[k * kNameBytes];
    const std::string name(s, strnlen(s, kNameBytes));
    f->ids[name] = (uint16_t)f->names.size();
    f->names.push_back(name);
  }
  f->names_written = f->names.size();
  std::vector<GsdEntry> ix(f->h.index_entries);
  if (!ix.empty() && !get(f, f->h.index_location, ix.data(), ix.size() * kEntryBytes)) return false;
  for (const GsdEntry& e : ix) {
    if (e.location == 0) break;  // the first unused entry ends the index
    if (e.id >= f->names.size() || type_bytes(e.type) == 0 ||
        e.location + e.rows * e.columns * type_bytes(e.type) > f->file_len)
      return false;
    f->index.push_back(e);
  }
  f->frames = f->index.empty() ? 0 : f->index.back().frame + 1;
  return true;
}

}  // namespace

extern "C" {

int hmc_gsd_create(const char* path, const char* application, const char* schema, uint32_t schema_major,
                   uint32_t schema_minor, hmc_gsd** out) {
  if (!path || !application || !schema || !out || schema_major > 0xffff || schema_minor > 0xffff)
    return HMC_ERR_INVALID;
  *out = nullptr;
  if (std::strlen(application) > kNameBytes - 1 || std::strlen(schema) > kNameBytes - 1) return HMC_ERR_INVALID;
  auto* f = new hmc_gsd;
  f->file = std::fopen(path, "w+b");
  if (!f->file) {
    delete f;
    return HMC_ERR_INVALID;
  }
  f->writable = true;
  f->h.magic = kGsdMagic;
  f->h.index_location = kHeaderBytes;
  f->h.index_entries = kFirstEntries;
  f->h.names_location = kHeaderBytes + kFirstEntries * kEntryBytes;
  f->h.names_entries = kFirstNames;
  f->h.schema_version = (schema_major << 16) | schema_minor;
  f->h.file_version = kFileVersion;
  std::memcpy(f->h.application, application, std::strlen(application));
  std::memcpy(f->h.schema, schema, std::strlen(schema));
  uint64_t at = 0;
  bool ok = put(f, 0, &f->h, kHeaderBytes);
  f->file_len = kHeaderBytes;
  ok = ok && grow(f, kFirstEntries * kEntryBytes + kFirstNames * kNameBytes, &at) && std::fflush(f->file) == 0;
  if (!ok) {
    std::fclose(f->file);
    delete f;
    return HMC_ERR_INVALID;
  }
  *out = f;
  return HMC_OK;
}

int hmc_gsd_open(const char* path, int writable, hmc_gsd** out) {
  if (!path || !out) return HMC_ERR_INVALID;
  *out = nullptr;
  auto* f = new hmc_gsd;
  f->file = std::fopen(path, writable ? "r+b" : "rb");
  f->writable = writable != 0;
  if (!f->file || !load(f)) {
    if (f->file) std::fclose(f->file);
    delete f;
    return HMC_ERR_INVALID;
  }
  *out = f;
  return HMC_OK;
}

int hmc_gsd_write_chunk(hmc_gsd* f, const char* name, uint8_t type, uint64_t rows, uint32_t columns,
                        const void* data) {
  if (!f || !name || !f->writable || columns == 0 || type_bytes(type) == 0 || (rows && !data))
    return HMC_ERR_INVALID;
  const size_t len = std::strlen(name);
  if (len == 0 || len > kNameBytes - 1) return HMC_ERR_INVALID;
  uint16_t id = 0;
  const auto it = f->ids.find(name);
  if (it != f->ids.end()) {
    id = it->second;
  } else {
    if (f->names.size() >= 0xffff) return HMC_ERR_INVALID;
    id = (uint16_t)f->names.size();
    f->ids[name] = id;
    f->names.push_back(name);
  }
  if (!f->frame_ids.insert(id).second) return HMC_ERR_INVALID;  // once per frame
  const uint64_t bytes = rows * columns * type_bytes(type);
  GsdEntry e{};
  e.frame = f->frames;
  e.rows = rows;
  e.location = f->file_len + f->data.size();
  e.columns = columns;
  e.id = id;
  e.type = type;
  f->pending.push_back(e);
  const auto* p = static_cast<const uint8_t*>(data);
  f->data.insert(f->data.end(), p, p + bytes);
  return HMC_OK;
}

int hmc_gsd_end_frame(hmc_gsd* f) {
  if (!f || !f->writable) return HMC_ERR_INVALID;
  if (!flush_frame(f)) return HMC_ERR_INVALID;
  f->frames += 1;
  f->frame_ids.clear();
  return HMC_OK;
}

int hmc_gsd_append_frame(hmc_gsd* f, hmc_ctx* c, uint64_t step) {
  if (!f || !c || !f->writable) return HMC_ERR_INVALID;
  if (!c->have_state) return fail(c, HMC_ERR_STATE, "no state");
  if (!f->pending.empty()) return fail(c, HMC_ERR_INVALID, "a GSD frame is already in progress");
  const Geometry& g = c->g;
  const uint32_t N = g.n_bodies;
  const bool oriented = g.model == HMC_CONVEX_POLYHEDRON;
  std::vector<float> pos(3 * (size_t)N), quat(oriented ? 4 * (size_t)N : 0);
  if (N) {
    HIPCHK(c, hipSetDevice(c->device));
    float* dpos = nullptr;
    float* dquat = nullptr;
    hipError_t e = hipMallocAsync((void**)&dpos, pos.size() * sizeof(float), c->stream);
    if (e == hipSuccess && oriented) e = hipMallocAsync((void**)&dquat, quat.size() * sizeof(float), c->stream);
    if (e == hipSuccess) {
      k_pack_frame<<<grid_for(N, 256), 256, 0, c->stream>>>(c->bodies, N, g.dim, dpos, dquat);
      e = hipGetLastError();
    }
    if (e == hipSuccess)
      e = hipMemcpyAsync(pos.data(), dpos, pos.size() * sizeof(float), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess && oriented)
      e = hipMemcpyAsync(quat.data(), dquat, quat.size() * sizeof(float), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (dpos) (void)hipFreeAsync(dpos, c->stream);
    if (dquat) (void)hipFreeAsync(dquat, c->stream);
    if (e != hipSuccess) return fail(c, HMC_ERR_DEVICE, std::string("GSD frame: ") + hipGetErrorString(e));
  }
  // Hypercuboid::to_gsd_box (hoomd-geometry/src/shape/cuboid.rs:357-390):
  // the edges, no tilt, Lz = 0 in 2D; written as FP32
  const float box[6] = {(float)g.L[0], (float)g.L[1], g.dim == 3 ? (float)g.L[2] : 0.0f, 0.0f, 0.0f, 0.0f};
  const uint8_t dims = (uint8_t)g.dim;
  const uint32_t n32 = N;
  int rc = hmc_gsd_write_chunk(f, "configuration/step", HMC_GSD_U64, 1, 1, &step);
  if (!rc) rc = hmc_gsd_write_chunk(f, "configuration/box", HMC_GSD_F32, 6, 1, box);
  if (!rc) rc = hmc_gsd_write_chunk(f, "configuration/dimensions", HMC_GSD_U8, 1, 1, &dims);
  if (!rc) rc = hmc_gsd_write_chunk(f, "particles/N", HMC_GSD_U32, 1, 1, &n32);
  if (!rc) rc = hmc_gsd_write_chunk(f, "particles/position", HMC_GSD_F32, N, 3, pos.data());
  if (!rc && oriented) rc = hmc_gsd_write_chunk(f, "particles/orientation", HMC_GSD_F32, N, 4, quat.data());
  if (!rc) rc = hmc_gsd_end_frame(f);
  if (rc) {
    f->pending.clear();
    f->data.clear();
    f->frame_ids.clear();
    return fail(c, rc, "GSD frame: writing the file failed");
  }
  return HMC_OK;
}

uint64_t hmc_gsd_n_frames(const hmc_gsd* f) { return f ? f->frames : 0; }

int hmc_gsd_read_chunk(const hmc_gsd* f, uint64_t frame, const char* name, void* out, uint64_t capacity_bytes,
                       uint64_t* rows, uint32_t* columns, uint8_t* type) {
  if (!f || !name) return HMC_ERR_INVALID;
  const auto it = f->ids.find(name);
  if (it == f->ids.end()) return HMC_ERR_INVALID;
  GsdEntry key{};
  key.frame = frame;
  key.id = it->second;
  const auto e = std::lower_bound(f->index.begin(), f->index.end(), key, entry_less);
  if (e == f->index.end() || e->frame != frame || e->id != key.id) return HMC_ERR_INVALID;
  if (rows) *rows = e->rows;
  if (columns) *columns = e->columns;
  if (type) *type = e->type;
  const uint64_t bytes = e->rows * e->columns * type_bytes(e->type);
  if (out && bytes && capacity_bytes >= bytes && !get(f, e->location, out, (size_t)bytes)) return HMC_ERR_INVALID;
  return HMC_OK;
}

int hmc_gsd_header(const hmc_gsd* f, char* application, char* schema, uint32_t* schema_major,
                   uint32_t* schema_minor) {
  if (!f) return HMC_ERR_INVALID;
  if (application) std::memcpy(application, f->h.application, 64);
  if (schema) std::memcpy(schema, f->h.schema, 64);
  if (schema_major) *schema_major = f->h.schema_version >> 16;
  if (schema_minor) *schema_minor = f->h.schema_version & 0xffffu;
  return HMC_OK;
}

int hmc_gsd_close(hmc_gsd* f) {
  if (!f) return HMC_ERR_INVALID;
  int rc = HMC_OK;
  if (f->writable && !f->pending.empty()) rc = hmc_gsd_end_frame(f);
  if (std::fclose(f->file) != 0) rc = HMC_ERR_INVALID;
  delete f;
  return rc;
}

}  // extern "C"
