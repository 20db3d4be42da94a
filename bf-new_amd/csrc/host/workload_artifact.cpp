// workload_artifact.cpp -- BFGRPH07 / BFQRY007 reader and writer (reference
// behaviour: src/io/workload_artifact.cpp:21-395; the byte layout is the
// reference's, the I/O is this file's own: the file is mapped, the
// fixed-size graph records are decoded and encoded in parallel, and writes
// go out in large blocks).
#include "bfnew/workload_artifact.hpp"

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <bit>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <thread>
#include <utility>

namespace bfnew {
namespace {

constexpr unsigned char kGraphMagic[8] = {'B', 'F', 'G', 'R', 'P', 'H', '0', '7'};
constexpr unsigned char kQueryMagic[8] = {'B', 'F', 'Q', 'R', 'Y', '0', '0', '7'};
constexpr std::uint64_t kGraphHeaderBytes = 24;  // magic, version, vertex count, edge count
constexpr std::uint64_t kVertexBytes = 11;
constexpr std::uint64_t kEdgeBytes = 28;

template <typename T>
void store(unsigned char* p, T v) {
  using U = std::make_unsigned_t<T>;
  U u = static_cast<U>(v);
  for (std::size_t i = 0; i < sizeof(T); ++i, u = static_cast<U>(u >> 8)) p[i] = static_cast<unsigned char>(u & 0xFFU);
}

template <typename T>
T load(const unsigned char* p) {
  using U = std::make_unsigned_t<T>;
  U u = 0;
  for (std::size_t i = 0; i < sizeof(T); ++i) u = static_cast<U>(u | (static_cast<U>(p[i]) << (8 * i)));
  return static_cast<T>(u);
}

// Runs f(begin, end) over [0, n) on up to 16 threads (one for small n).
template <typename F>
void parallel_ranges(std::uint64_t n, F&& f) {
  const unsigned hw = std::max(1U, std::min(16U, std::thread::hardware_concurrency()));
  const std::uint64_t workers = std::max<std::uint64_t>(1U, std::min<std::uint64_t>(hw, n / 65536U));
  if (workers == 1U) {
    f(std::uint64_t{0}, n);
    return;
  }
  std::vector<std::thread> pool;
  for (std::uint64_t w = 0; w < workers; ++w)
    pool.emplace_back([&, w] { f(n * w / workers, n * (w + 1) / workers); });
  for (std::thread& t : pool) t.join();
}

// A read-only mapping of a whole file.
class MappedFile {
 public:
  explicit MappedFile(const std::filesystem::path& path) {
    fd_ = ::open(path.c_str(), O_RDONLY | O_CLOEXEC);
    if (fd_ < 0) throw std::runtime_error{"cannot open workload artifact: " + path.string()};
    struct stat st {};
    if (::fstat(fd_, &st) != 0) {
      ::close(fd_);
      throw std::runtime_error{"cannot open workload artifact: " + path.string()};
    }
    size_ = static_cast<std::uint64_t>(st.st_size);
    if (size_ > 0U) {
      void* p = ::mmap(nullptr, size_, PROT_READ, MAP_PRIVATE, fd_, 0);
      if (p == MAP_FAILED) {
        ::close(fd_);
        throw std::runtime_error{"cannot map workload artifact: " + path.string()};
      }
      ::madvise(p, size_, MADV_SEQUENTIAL);
      data_ = static_cast<const unsigned char*>(p);
    }
  }
  ~MappedFile() {
    if (data_) ::munmap(const_cast<unsigned char*>(data_), size_);
    if (fd_ >= 0) ::close(fd_);
  }
  MappedFile(const MappedFile&) = delete;
  MappedFile& operator=(const MappedFile&) = delete;
  [[nodiscard]] const unsigned char* data() const noexcept { return data_; }
  [[nodiscard]] std::uint64_t size() const noexcept { return size_; }

 private:
  int fd_ = -1;
  const unsigned char* data_ = nullptr;
  std::uint64_t size_ = 0;
};

// Sequential decoding of the variable-length query records.
class Cursor {
 public:
  Cursor(const unsigned char* p, std::uint64_t n) : p_{p}, left_{n} {}
  const unsigned char* take(std::uint64_t n) {
    if (n > left_) throw std::runtime_error{"truncated workload artifact"};
    const unsigned char* at = p_;
    p_ += n;
    left_ -= n;
    return at;
  }
  template <typename T>
  T value() {
    return load<T>(take(sizeof(T)));
  }
  bool boolean() {
    const std::uint8_t b = value<std::uint8_t>();
    if (b > 1U) throw std::runtime_error{"invalid artifact boolean"};
    return b != 0U;
  }
  std::string text() {
    const std::uint32_t n = value<std::uint32_t>();
    const unsigned char* s = take(n);
    return std::string(reinterpret_cast<const char*>(s), n);
  }
  template <typename Id>
  std::vector<Id> ids() {
    const std::uint32_t n = value<std::uint32_t>();
    const unsigned char* s = take(static_cast<std::uint64_t>(n) * sizeof(typename Id::representation_type));
    std::vector<Id> out(n);
    for (std::uint32_t i = 0; i < n; ++i)
      out[i] = Id{load<typename Id::representation_type>(s + i * sizeof(typename Id::representation_type))};
    return out;
  }
  std::vector<std::uint32_t> words() {
    const std::uint32_t n = value<std::uint32_t>();
    const unsigned char* s = take(4ULL * n);
    std::vector<std::uint32_t> out(n);
    for (std::uint32_t i = 0; i < n; ++i) out[i] = load<std::uint32_t>(s + 4ULL * i);
    return out;
  }
  void magic(const unsigned char (&expected)[8]) {
    if (std::memcmp(take(8), expected, 8) != 0) throw std::runtime_error{"unrecognized workload artifact magic"};
  }
  [[nodiscard]] std::uint64_t left() const noexcept { return left_; }

 private:
  const unsigned char* p_;
  std::uint64_t left_;
};

// Buffered little-endian encoding, written out in large blocks.
class Writer {
 public:
  explicit Writer(const std::filesystem::path& path) {
    if (const auto parent = path.parent_path(); !parent.empty()) std::filesystem::create_directories(parent);
    fd_ = ::open(path.c_str(), O_WRONLY | O_CREAT | O_TRUNC | O_CLOEXEC, 0644);
    if (fd_ < 0) throw std::runtime_error{"cannot create workload artifact: " + path.string()};
  }
  ~Writer() {
    if (fd_ >= 0) ::close(fd_);
  }
  Writer(const Writer&) = delete;
  Writer& operator=(const Writer&) = delete;

  std::vector<unsigned char>& buffer() noexcept { return buf_; }
  template <typename T>
  void value(T v) {
    const std::size_t at = buf_.size();
    buf_.resize(at + sizeof(T));
    store(buf_.data() + at, v);
  }
  void bytes(const void* p, std::size_t n) {
    const auto* c = static_cast<const unsigned char*>(p);
    buf_.insert(buf_.end(), c, c + n);
  }
  void boolean(bool b) { value<std::uint8_t>(b ? 1U : 0U); }
  void text(const std::string& s) {
    value(checked_length(s.size(), "artifact string exceeds 32-bit length"));
    bytes(s.data(), s.size());
  }
  template <typename Id>
  void ids(const std::vector<Id>& v) {
    value(checked_length(v.size(), "artifact ID vector exceeds 32-bit length"));
    for (const Id& x : v) value(x.value());
  }
  void words(const std::vector<std::uint32_t>& v) {
    value(checked_length(v.size(), "artifact index vector exceeds 32-bit length"));
    for (std::uint32_t x : v) value(x);
  }
  void flush() {
    const unsigned char* p = buf_.data();
    std::size_t n = buf_.size();
    while (n > 0U) {
      const ssize_t w = ::write(fd_, p, std::min<std::size_t>(n, std::size_t{1} << 30));
      if (w <= 0) throw std::runtime_error{"failed while writing workload artifact"};
      p += w;
      n -= static_cast<std::size_t>(w);
    }
    buf_.clear();
  }
  void close() {
    flush();
    if (::close(fd_) != 0) {
      fd_ = -1;
      throw std::runtime_error{"failed while writing workload artifact"};
    }
    fd_ = -1;
  }

 private:
  static std::uint32_t checked_length(std::size_t n, const char* what) {
    if (n > std::numeric_limits<std::uint32_t>::max()) throw std::length_error{what};
    return static_cast<std::uint32_t>(n);
  }
  int fd_ = -1;
  std::vector<unsigned char> buf_;
};

void write_terminal(Writer& w, const MappedTerminal& t) {
  w.text(t.site);
  w.text(t.pin);
  w.value(t.original_vertex.value());
  w.value(t.vertex.value());
  w.value(t.x);
  w.value(t.y);
  w.boolean(t.has_location);
  w.value(t.owner_tile.value());
}

MappedTerminal read_terminal(Cursor& c) {
  MappedTerminal t;
  t.site = c.text();
  t.pin = c.text();
  t.original_vertex = VertexId{c.value<std::uint32_t>()};
  t.vertex = VertexId{c.value<std::uint32_t>()};
  t.x = c.value<std::int32_t>();
  t.y = c.value<std::int32_t>();
  t.has_location = c.boolean();
  t.owner_tile = TileId{c.value<std::uint32_t>()};
  return t;
}

void check_terminal(const MappedTerminal& t, const WeightedGraph& g) {
  const auto to_new = g.old_to_new();
  if (t.vertex.value() >= g.vertex_count() || t.original_vertex.value() >= to_new.size() ||
      to_new[t.original_vertex.value()] != t.vertex)
    throw std::runtime_error{"query artifact terminal vertex mapping is invalid"};
  const VertexMetadata& m = g.vertices()[t.vertex.value()];
  if (m.x != t.x || m.y != t.y || m.has_location != t.has_location || g.owner_tiles()[t.vertex.value()] != t.owner_tile)
    throw std::runtime_error{"query artifact terminal spatial metadata is invalid"};
}

}  // namespace

void write_graph_artifact(const std::filesystem::path& path, const InputGraph& graph) {
  Writer w{path};
  w.bytes(kGraphMagic, 8);
  w.value(workload_artifact_version);
  w.value(graph.vertex_count());
  w.value(static_cast<std::uint64_t>(graph.edge_count()));
  const auto vertices = graph.vertices();
  const auto edges = graph.edges();
  // Vertices, then edges in blocks of 1 << 22 records, each block encoded on
  // several threads before it is written.
  auto& buf = w.buffer();
  const std::size_t vbase = buf.size();
  buf.resize(vbase + kVertexBytes * vertices.size());
  parallel_ranges(vertices.size(), [&](std::uint64_t a, std::uint64_t b) {
    for (std::uint64_t i = a; i < b; ++i) {
      unsigned char* p = buf.data() + vbase + kVertexBytes * i;
      store(p, vertices[i].x);
      store(p + 4, vertices[i].y);
      p[8] = vertices[i].has_location ? 1U : 0U;
      store(p + 9, vertices[i].resource_class.value());
    }
  });
  w.flush();
  constexpr std::uint64_t kBlock = std::uint64_t{1} << 22;
  for (std::uint64_t e0 = 0; e0 < edges.size(); e0 += kBlock) {
    const std::uint64_t n = std::min<std::uint64_t>(kBlock, edges.size() - e0);
    buf.resize(kEdgeBytes * n);
    parallel_ranges(n, [&](std::uint64_t a, std::uint64_t b) {
      for (std::uint64_t i = a; i < b; ++i) {
        const EdgeInputRecord& e = edges[e0 + i];
        unsigned char* p = buf.data() + kEdgeBytes * i;
        store(p, e.source.value());
        store(p + 4, e.destination.value());
        store(p + 8, canonical_weight_bits(e.weight));
        store(p + 12, e.provenance.domain);
        store(p + 16, e.provenance.kind_and_flags);
        store(p + 20, e.provenance.source_record);
      }
    });
    w.flush();
  }
  w.close();
}

InputGraph read_graph_artifact(const std::filesystem::path& path) {
  const MappedFile file{path};
  Cursor head{file.data(), file.size()};
  head.magic(kGraphMagic);
  if (head.value<std::uint32_t>() != workload_artifact_version)
    throw std::runtime_error{"unsupported graph artifact version"};
  const std::uint32_t vertex_count = head.value<std::uint32_t>();
  const std::uint64_t edge_count = head.value<std::uint64_t>();
  const unsigned __int128 expected = static_cast<unsigned __int128>(kGraphHeaderBytes) +
                                     static_cast<unsigned __int128>(kVertexBytes) * vertex_count +
                                     static_cast<unsigned __int128>(kEdgeBytes) * edge_count;
  if (expected > file.size()) throw std::runtime_error{"truncated workload artifact"};
  if (expected < file.size()) throw std::runtime_error{"workload artifact has trailing data"};
  if (edge_count > std::numeric_limits<std::size_t>::max() / sizeof(EdgeInputRecord))
    throw std::length_error{"graph artifact edge count exceeds host size"};

  const unsigned char* vbase = file.data() + kGraphHeaderBytes;
  std::vector<VertexMetadata> vertices(vertex_count);
  std::atomic<bool> bad_boolean{false};
  parallel_ranges(vertex_count, [&](std::uint64_t a, std::uint64_t b) {
    for (std::uint64_t i = a; i < b; ++i) {
      const unsigned char* p = vbase + kVertexBytes * i;
      if (p[8] > 1U) bad_boolean.store(true, std::memory_order_relaxed);
      vertices[i] = VertexMetadata{load<std::int32_t>(p), load<std::int32_t>(p + 4), p[8] == 1U,
                                   ResourceClassId{load<std::uint16_t>(p + 9)}};
    }
  });
  if (bad_boolean.load()) throw std::runtime_error{"invalid artifact boolean"};
  const unsigned char* ebase = vbase + kVertexBytes * vertex_count;
  std::vector<EdgeInputRecord> edges(static_cast<std::size_t>(edge_count));
  parallel_ranges(edge_count, [&](std::uint64_t a, std::uint64_t b) {
    for (std::uint64_t i = a; i < b; ++i) {
      const unsigned char* p = ebase + kEdgeBytes * i;
      EdgeInputRecord& e = edges[i];
      e.source = VertexId{load<std::uint32_t>(p)};
      e.destination = VertexId{load<std::uint32_t>(p + 4)};
      e.weight = std::bit_cast<float>(load<std::uint32_t>(p + 8));
      e.provenance = PhysicalProvenance{load<std::uint32_t>(p + 12), load<std::uint32_t>(p + 16),
                                        load<std::uint64_t>(p + 20)};
    }
  });
  return InputGraph(std::move(vertices), std::move(edges));
}

void write_query_artifact(const std::filesystem::path& path, const std::vector<MappedRouteQuery>& queries) {
  Writer w{path};
  w.bytes(kQueryMagic, 8);
  w.value(workload_artifact_version);
  w.value(static_cast<std::uint64_t>(queries.size()));
  for (const MappedRouteQuery& m : queries) {
    const RouteQuery& q = m.query;
    w.text(m.net_name);
    w.value(q.query_id.value());
    w.ids(q.source_terminals);
    w.ids(q.target_terminals);
    w.ids(q.sources);
    w.ids(q.targets);
    w.words(q.source_terminal_to_source);
    w.words(q.target_terminal_to_target);
    w.ids(q.selected_tiles);
    w.value(q.expansion_generation);
    if (m.source_terminals.size() > std::numeric_limits<std::uint32_t>::max() ||
        m.target_terminals.size() > std::numeric_limits<std::uint32_t>::max())
      throw std::length_error{"artifact terminal vector exceeds 32-bit length"};
    w.value(static_cast<std::uint32_t>(m.source_terminals.size()));
    for (const MappedTerminal& t : m.source_terminals) write_terminal(w, t);
    w.value(static_cast<std::uint32_t>(m.target_terminals.size()));
    for (const MappedTerminal& t : m.target_terminals) write_terminal(w, t);
    if (w.buffer().size() >= (std::size_t{1} << 24)) w.flush();
  }
  w.close();
}

std::vector<MappedRouteQuery> read_query_artifact(const std::filesystem::path& path,
                                                  const WeightedGraph& spatial_graph) {
  const MappedFile file{path};
  Cursor c{file.data(), file.size()};
  c.magic(kQueryMagic);
  if (c.value<std::uint32_t>() != workload_artifact_version)
    throw std::runtime_error{"unsupported query artifact version"};
  const std::uint64_t count = c.value<std::uint64_t>();
  std::vector<MappedRouteQuery> out;
  out.reserve(static_cast<std::size_t>(std::min<std::uint64_t>(count, c.left() / 64U + 1U)));
  for (std::uint64_t i = 0; i < count; ++i) {
    MappedRouteQuery m;
    RouteQuery& q = m.query;
    m.net_name = c.text();
    q.query_id = QueryId{c.value<std::uint32_t>()};
    q.source_terminals = c.ids<VertexId>();
    q.target_terminals = c.ids<VertexId>();
    q.sources = c.ids<VertexId>();
    q.targets = c.ids<VertexId>();
    q.source_terminal_to_source = c.words();
    q.target_terminal_to_target = c.words();
    q.selected_tiles = c.ids<TileId>();
    q.expansion_generation = c.value<std::uint32_t>();
    const std::uint32_t ns = c.value<std::uint32_t>();
    for (std::uint32_t k = 0; k < ns; ++k) m.source_terminals.push_back(read_terminal(c));
    const std::uint32_t nt = c.value<std::uint32_t>();
    for (std::uint32_t k = 0; k < nt; ++k) m.target_terminals.push_back(read_terminal(c));
    if (m.source_terminals.size() != q.source_terminals.size() ||
        m.target_terminals.size() != q.target_terminals.size())
      throw std::runtime_error{"query artifact terminal descriptor count mismatch"};
    if (!validate_route_query(spatial_graph, q).ok())
      throw std::runtime_error{"query artifact contains an invalid RouteQuery"};
    for (std::size_t k = 0; k < m.source_terminals.size(); ++k) {
      check_terminal(m.source_terminals[k], spatial_graph);
      if (m.source_terminals[k].vertex != q.source_terminals[k])
        throw std::runtime_error{"source descriptor and query terminal disagree"};
    }
    for (std::size_t k = 0; k < m.target_terminals.size(); ++k) {
      check_terminal(m.target_terminals[k], spatial_graph);
      if (m.target_terminals[k].vertex != q.target_terminals[k])
        throw std::runtime_error{"target descriptor and query terminal disagree"};
    }
    out.push_back(std::move(m));
  }
  if (c.left() != 0U) throw std::runtime_error{"workload artifact has trailing data"};
  return out;
}

bool files_are_byte_identical(const std::filesystem::path& left, const std::filesystem::path& right) {
  const MappedFile a{left};
  const MappedFile b{right};
  return a.size() == b.size() && (a.size() == 0U || std::memcmp(a.data(), b.data(), a.size()) == 0);
}

}  // namespace bfnew
