// svo_octree_tool — reader, validator and diff for .octree / .octreenodes / .octreedata
// (the formats of octree_io.h:49-110, Node.h:16-65, VoxelData.h:10-18).
//
//   svo_octree_tool info  <base>.octree
//   svo_octree_tool check <base>.octree
//   svo_octree_tool diff  <a>.octree <b>.octree
//
// The node and data files are memory-mapped, so outputs of tens of GB are checked without
// loading them. 'check' verifies everything the streaming OctreeBuilder
// (OctreeBuilder.cpp:34-168) guarantees about its output:
//   * header fields and file sizes (24-byte nodes, 32-byte data records);
//   * data[0] is the null record; a geometry-only build's data[1] is the white record;
//   * the tree below the root (the last node) reaches every other node exactly once, with
//     child offsets that are the ranks of the present children in child order and leaves
//     exactly at depth log2(gridlength);
//   * the file order: walking the internal nodes in post-order, each one's children are
//     the next records of the file (groupNodes writes them when the node completes);
//   * leaf codes from the walk are strictly increasing; for payload builds every leaf's
//     record carries its own Morton code; data indices follow the reference's order
//     (leaves in Morton order without -levels; post-order over all nodes with -levels).
// Exit status 0 = valid / identical, 1 = invalid / different, 2 = usage or I/O error.
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

namespace {

struct Header {
  int version = 0;
  uint64_t gridlength = 0, n_nodes = 0, n_data = 0;
};

// octree_io.h:86-110: "#octreeheader <v>", then keywords until END; unknown keywords are
// skipped to the end of their line.
bool parse_header(const std::string& path, Header& h, std::string& err) {
  std::ifstream f(path);
  if (!f) {
    err = "cannot open " + path;
    return false;
  }
  std::string line;
  f >> line >> h.version;
  if (line != "#octreeheader") {
    err = "first line reads [" + line + "] instead of [#octreeheader]";
    return false;
  }
  bool done = false;
  while (f.good() && !done) {
    f >> line;
    if (line == "END")
      done = true;
    else if (line == "gridlength")
      f >> h.gridlength;
    else if (line == "n_nodes")
      f >> h.n_nodes;
    else if (line == "n_data")
      f >> h.n_data;
    else {
      char c;
      do {
        c = (char)f.get();
      } while (f.good() && c != '\n');
    }
  }
  if (!done) {
    err = "header has no END";
    return false;
  }
  return true;
}

struct Mapped {
  const uint8_t* p = nullptr;
  size_t size = 0;
  bool open(const std::string& path, std::string& err) {
    int fd = ::open(path.c_str(), O_RDONLY);
    if (fd < 0) {
      err = "cannot open " + path;
      return false;
    }
    struct stat st;
    if (fstat(fd, &st) != 0) {
      ::close(fd);
      err = "cannot stat " + path;
      return false;
    }
    size = (size_t)st.st_size;
    if (size) {
      void* m = mmap(nullptr, size, PROT_READ, MAP_PRIVATE, fd, 0);
      if (m == MAP_FAILED) {
        ::close(fd);
        err = "cannot map " + path;
        return false;
      }
      p = (const uint8_t*)m;
    }
    ::close(fd);
    return true;
  }
  ~Mapped() {
    if (p) munmap((void*)p, size);
  }
};

struct NodeRec {
  uint64_t data, children_base;
  int8_t off[8];
};
static_assert(sizeof(NodeRec) == 24, "node record is 24 bytes");

struct DataRec {
  uint64_t morton;
  float color[3], normal[3];
};
static_assert(sizeof(DataRec) == 32, "data record is 32 bytes");

std::string base_of(const std::string& header_path) {
  const size_t dot = header_path.find_last_of('.');
  return dot == std::string::npos ? header_path : header_path.substr(0, dot);
}

struct Octree {
  Header h;
  Mapped nodes, data;
  const NodeRec* node(uint64_t i) const { return (const NodeRec*)nodes.p + i; }
  const DataRec* rec(uint64_t i) const { return (const DataRec*)data.p + i; }
  bool load(const std::string& header_path, std::string& err) {
    if (!parse_header(header_path, h, err)) return false;
    const std::string b = base_of(header_path);
    return nodes.open(b + ".octreenodes", err) && data.open(b + ".octreedata", err);
  }
};

bool is_leaf(const NodeRec& n) {
  for (int k = 0; k < 8; k++)
    if (n.off[k] != -1) return false;
  return true;
}

int log2_exact(uint64_t g) {
  int d = 0;
  while ((1ull << d) < g) d++;
  return ((1ull << d) == g) ? d : -1;
}

struct Report {
  uint64_t leaves = 0, internal = 0, max_depth = 0;
  bool payload = false, levels = false;
};

#define FAIL(msg)            \
  do {                       \
    err = (msg);             \
    return false;            \
  } while (0)

bool check(const Octree& t, Report& rep, std::string& err) {
  const Header& h = t.h;
  const int D = log2_exact(h.gridlength);
  if (h.version != 1) FAIL("header version is not 1");
  if (D < 0) FAIL("gridlength is not a power of 2");
  if (h.n_nodes == 0) FAIL("n_nodes is 0");
  if (t.nodes.size != h.n_nodes * 24)
    FAIL("node file holds " + std::to_string(t.nodes.size) + " bytes, header says " +
         std::to_string(h.n_nodes) + " nodes");
  if (t.data.size != h.n_data * 32)
    FAIL("data file holds " + std::to_string(t.data.size) + " bytes, header says " +
         std::to_string(h.n_data) + " records");
  if (h.n_data < 1) FAIL("no null data record");
  static const uint8_t zeros[32] = {0};
  if (memcmp(t.rec(0), zeros, 32) != 0) FAIL("data[0] is not the null record");

  // Walk from the root. Post-order over internal nodes gives the expected position of each
  // node's children group; leaves are visited in Morton order.
  struct Frame {
    uint64_t idx, prefix;
    int depth, next_k;
  };
  const uint64_t root = h.n_nodes - 1;
  std::vector<uint8_t> seen(h.n_nodes, 0);
  std::vector<Frame> st;
  st.push_back({root, 0, 0, 0});
  seen[root] = 1;
  uint64_t next_group = 0;   // the next children group must start here
  uint64_t post_rank = 0;    // rank among nodes that write data (post-order)
  uint64_t leaf_rank = 0;
  uint64_t last_code = 0;
  bool have_code = false;
  // -levels: internal nodes carry data, the root included (it is internal unless the
  // tree is a single leaf or the null root)
  const int levels = (h.n_nodes > 1 && t.node(root)->data != 0) ? 1 : 0;
  // payload or geometry only, from the first two leaves in Morton order: geometry-only
  // leaves all point at record 1 (the white record), payload leaves at their own records
  int mode = -1;  // 0 binary, 1 payload, -1 no leaves
  {
    uint64_t found = 0, d1 = 0, d2 = 0;
    std::vector<Frame> w;
    w.push_back({root, 0, 0, 0});
    while (!w.empty() && found < 2) {
      Frame& f = w.back();
      const NodeRec& n = *t.node(f.idx);
      if (f.depth == D || is_leaf(n)) {
        if (f.depth == D && n.data) {
          (found == 0 ? d1 : d2) = n.data;
          found++;
        }
        w.pop_back();
        continue;
      }
      int k = f.next_k;
      while (k < 8 && n.off[k] == -1) k++;
      if (k == 8 || n.children_base + (uint64_t)n.off[k] >= root) {
        w.pop_back();
        continue;
      }
      f.next_k = k + 1;
      w.push_back({n.children_base + (uint64_t)n.off[k], (f.prefix << 3) | (uint64_t)k,
                   f.depth + 1, 0});
    }
    if (found == 2) {
      mode = (d2 == 1) ? 0 : 1;
    } else if (found == 1) {
      const DataRec* r = h.n_data > d1 ? t.rec(d1) : nullptr;
      const bool white = r && r->morton == 0 && r->color[0] == 1.0f && r->color[1] == 1.0f &&
                         r->color[2] == 1.0f && r->normal[0] == 0.0f && r->normal[1] == 0.0f &&
                         r->normal[2] == 0.0f;
      // a single voxel: geometry only when its record is the white one (a payload voxel at
      // code 0 with exactly that colour and normal reads the same either way)
      mode = (d1 == 1 && white) ? 0 : 1;
    }
  }
  while (!st.empty()) {
    Frame& f = st.back();
    const NodeRec& n = *t.node(f.idx);
    if (f.depth == D || is_leaf(n)) {
      // the null root of an empty tree (finalizeTree pads the whole grid)
      if (f.idx == root && h.n_nodes == 1 && n.data == 0 && n.children_base == 0 && is_leaf(n)) {
        st.pop_back();
        continue;
      }
      if (f.depth != D)
        FAIL("leaf record " + std::to_string(f.idx) + " at depth " + std::to_string(f.depth) +
             " (leaves live at depth " + std::to_string(D) + ")");
      if (!is_leaf(n)) FAIL("node " + std::to_string(f.idx) + " below the leaves");
      {
        if (n.children_base != 0) FAIL("leaf " + std::to_string(f.idx) + " has children_base");
        if (n.data == 0) FAIL("leaf " + std::to_string(f.idx) + " is null");
        if (have_code && f.prefix <= last_code) FAIL("leaf codes not increasing");
        last_code = f.prefix;
        have_code = true;
        if (mode == 0) {
          if (n.data != 1) FAIL("geometry-only leaf " + std::to_string(f.idx) + " data != 1");
        } else {
          if (n.data >= h.n_data) FAIL("leaf data index out of range");
          if (t.rec(n.data)->morton != f.prefix)
            FAIL("leaf " + std::to_string(f.idx) + ": its record carries morton " +
                 std::to_string(t.rec(n.data)->morton) + ", its code is " +
                 std::to_string(f.prefix));
          // data order: Morton rank without -levels, post-order rank with -levels
          const uint64_t want = levels ? 1 + post_rank : 1 + leaf_rank;
          if (n.data != want)
            FAIL("leaf " + std::to_string(f.idx) + " data index " + std::to_string(n.data) +
                 ", expected " + std::to_string(want));
          post_rank++;
        }
        leaf_rank++;
      }
      rep.max_depth = std::max<uint64_t>(rep.max_depth, (uint64_t)f.depth);
      st.pop_back();
      continue;
    }
    if (f.next_k == 0) {
      // first visit: offsets are the ranks of the present children
      int r = 0;
      for (int k = 0; k < 8; k++) {
        if (n.off[k] == -1) continue;
        if (n.off[k] != r) FAIL("node " + std::to_string(f.idx) + ": offset[" + std::to_string(k) +
                                "] = " + std::to_string(n.off[k]) + ", expected " +
                                std::to_string(r));
        r++;
      }
      if (n.children_base + (uint64_t)r > root)
        FAIL("node " + std::to_string(f.idx) + ": children beyond the root");
    }
    // descend into the next present child
    int k = f.next_k;
    while (k < 8 && n.off[k] == -1) k++;
    if (k < 8) {
      f.next_k = k + 1;
      const uint64_t c = n.children_base + (uint64_t)n.off[k];
      if (c >= root || seen[c]) FAIL("node " + std::to_string(c) + " reached twice or is the root");
      seen[c] = 1;
      const Frame child{c, (f.prefix << 3) | (uint64_t)k, f.depth + 1, 0};
      st.push_back(child);
      continue;
    }
    // all children done: this internal node completes (post-order)
    int cnt = 0;
    for (int q = 0; q < 8; q++) cnt += n.off[q] != -1;
    if (n.children_base != next_group)
      FAIL("node " + std::to_string(f.idx) + ": children at " + std::to_string(n.children_base) +
           ", the streaming order puts them at " + std::to_string(next_group));
    next_group += (uint64_t)cnt;
    rep.internal++;
    // data: 0 without -levels; a post-order index with -levels
    if (levels) {
      const uint64_t want = (mode == 1) ? 1 + post_rank : 2 + post_rank;
      if (n.data != want)
        FAIL("internal node " + std::to_string(f.idx) + " data index " + std::to_string(n.data) +
             ", expected " + std::to_string(want));
      if (want >= h.n_data || t.rec(n.data)->morton != 0)
        FAIL("internal node " + std::to_string(f.idx) + ": no averaged record at its index");
      post_rank++;
    } else if (n.data != 0) {
      FAIL("internal node " + std::to_string(f.idx) + " has data without -levels");
    }
    st.pop_back();
  }
  if (h.n_nodes > 1 && next_group != root)
    FAIL("children groups end at " + std::to_string(next_group) + ", the root is " +
         std::to_string(root));
  for (uint64_t i = 0; i < h.n_nodes; i++)
    if (!seen[i]) FAIL("node " + std::to_string(i) + " is not reachable from the root");
  rep.leaves = leaf_rank;
  rep.payload = mode == 1;
  rep.levels = levels == 1;
  const uint64_t want_data = (rep.payload ? 1 + rep.leaves : 2) + (rep.levels ? rep.internal : 0);
  if (mode == -1) {
    // an empty tree: binary (2 records) or payload (1 record)
    if (h.n_data != 1 && h.n_data != 2) FAIL("empty tree with " + std::to_string(h.n_data) + " records");
  } else if (h.n_data != want_data) {
    FAIL("n_data " + std::to_string(h.n_data) + ", the tree needs " + std::to_string(want_data));
  }
  if (mode == 0 || (mode == -1 && h.n_data == 2)) {
    const DataRec* w = t.rec(1);
    if (w->morton != 0 || w->color[0] != 1.0f || w->color[1] != 1.0f || w->color[2] != 1.0f ||
        w->normal[0] != 0.0f || w->normal[1] != 0.0f || w->normal[2] != 0.0f)
      FAIL("data[1] is not the white record of a geometry-only build");
  }
  return true;
}

int cmd_info(const std::string& path) {
  Header h;
  std::string err;
  if (!parse_header(path, h, err)) {
    std::cerr << "Error: " << err << std::endl;
    return 2;
  }
  std::cout << "  version: " << h.version << std::endl
            << "  base_filename: " << base_of(path) << std::endl
            << "  grid length: " << h.gridlength << std::endl
            << "  n_nodes: " << h.n_nodes << std::endl
            << "  n_data: " << h.n_data << std::endl;
  return 0;
}

int cmd_check(const std::string& path) {
  Octree t;
  std::string err;
  if (!t.load(path, err)) {
    std::cerr << "Error: " << err << std::endl;
    return 2;
  }
  Report rep;
  if (!check(t, rep, err)) {
    std::cout << "INVALID: " << err << std::endl;
    return 1;
  }
  std::cout << "OK: gridlength " << t.h.gridlength << ", " << t.h.n_nodes << " nodes ("
            << rep.internal << " internal, " << rep.leaves << " leaves), " << t.h.n_data
            << " data records, " << (rep.payload ? "payload" : "geometry only")
            << (rep.levels ? ", with levels" : "") << std::endl;
  return 0;
}

int cmd_diff(const std::string& a, const std::string& b) {
  Octree x, y;
  std::string err;
  if (!x.load(a, err) || !y.load(b, err)) {
    std::cerr << "Error: " << err << std::endl;
    return 2;
  }
  if (x.h.gridlength != y.h.gridlength || x.h.n_nodes != y.h.n_nodes || x.h.n_data != y.h.n_data) {
    std::cout << "DIFFERENT headers: gridlength " << x.h.gridlength << "/" << y.h.gridlength
              << ", n_nodes " << x.h.n_nodes << "/" << y.h.n_nodes << ", n_data " << x.h.n_data
              << "/" << y.h.n_data << std::endl;
    return 1;
  }
  const uint64_t nn = std::min<uint64_t>(x.nodes.size, y.nodes.size) / 24;
  for (uint64_t i = 0; i < nn; i++) {
    if (memcmp(x.node(i), y.node(i), 24) != 0) {
      const NodeRec &p = *x.node(i), &q = *y.node(i);
      std::cout << "DIFFERENT node " << i << ": data " << p.data << "/" << q.data
                << ", children_base " << p.children_base << "/" << q.children_base
                << ", offsets";
      for (int k = 0; k < 8; k++) std::cout << " " << (int)p.off[k] << "/" << (int)q.off[k];
      std::cout << std::endl;
      return 1;
    }
  }
  const uint64_t nd = std::min<uint64_t>(x.data.size, y.data.size) / 32;
  for (uint64_t i = 0; i < nd; i++) {
    if (memcmp(x.rec(i), y.rec(i), 32) != 0) {
      const DataRec &p = *x.rec(i), &q = *y.rec(i);
      std::cout << "DIFFERENT data " << i << ": morton " << p.morton << "/" << q.morton;
      for (int k = 0; k < 3; k++) std::cout << ", color[" << k << "] " << p.color[k] << "/" << q.color[k];
      for (int k = 0; k < 3; k++)
        std::cout << ", normal[" << k << "] " << p.normal[k] << "/" << q.normal[k];
      std::cout << std::endl;
      return 1;
    }
  }
  if (x.nodes.size != y.nodes.size || x.data.size != y.data.size) {
    std::cout << "DIFFERENT file sizes" << std::endl;
    return 1;
  }
  std::cout << "IDENTICAL" << std::endl;
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  const std::string usage =
      "usage: svo_octree_tool info <f.octree> | check <f.octree> | diff <a.octree> <b.octree>";
  if (argc < 3) {
    std::cerr << usage << std::endl;
    return 2;
  }
  const std::string cmd = argv[1];
  if (cmd == "info" && argc == 3) return cmd_info(argv[2]);
  if (cmd == "check" && argc == 3) return cmd_check(argv[2]);
  if (cmd == "diff" && argc == 4) return cmd_diff(argv[2], argv[3]);
  std::cerr << usage << std::endl;
  return 2;
}
