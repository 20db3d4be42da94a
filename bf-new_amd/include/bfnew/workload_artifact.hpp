// workload_artifact.hpp -- the BFGRPH07 graph and BFQRY007 query artifact
// files (reference: src/io/workload_artifact.cpp:21-395,
// include/bfnew/fpga_interchange.hpp:15, 83-103, 128-150).
//
// Both are little-endian with no padding.  Graph: "BFGRPH07", u32 version
// (1), u32 vertex count, u64 edge count; per vertex i32 x, i32 y, u8
// has_location (0/1), u16 resource class (11 bytes); per edge u32 source,
// u32 destination, u32 canonical weight bits (-0 written as +0), u32
// provenance domain, u32 kind_and_flags, u64 source record (28 bytes).
// Queries: "BFQRY007", u32 version, u64 query count; per query a string net
// name (u32 length + bytes), u32 query id, the id vectors source_terminals,
// target_terminals, sources, targets (u32 count + u32 each), the u32
// vectors source_terminal_to_source and target_terminal_to_target, the tile
// vector selected_tiles, u32 expansion generation, then u32 count + source
// terminal descriptors and u32 count + target terminal descriptors (site and
// pin strings, u32 original vertex, u32 vertex, i32 x, i32 y, u8
// has_location, u32 owner tile).  A file must end exactly after its records.
//
// Reading maps the file and decodes the fixed-size graph records on several
// threads; writing encodes them the same way and writes large blocks.
#pragma once

#include "bfnew/graph.hpp"
#include "bfnew/query.hpp"

#include <cstdint>
#include <filesystem>
#include <string>
#include <vector>

namespace bfnew {

inline constexpr std::uint32_t workload_artifact_version = 1U;

struct MappedTerminal {
  std::string site;
  std::string pin;
  VertexId original_vertex{};
  VertexId vertex{};
  std::int32_t x{};
  std::int32_t y{};
  bool has_location{};
  TileId owner_tile{};
  bool operator==(const MappedTerminal&) const = default;
};

struct MappedRouteQuery {
  std::string net_name;
  RouteQuery query;
  std::vector<MappedTerminal> source_terminals;
  std::vector<MappedTerminal> target_terminals;
  bool operator==(const MappedRouteQuery&) const = default;
};

// Creates parent directories.  Throws std::runtime_error on I/O failure and
// std::length_error for a string or vector longer than 2^32 - 1.
void write_graph_artifact(const std::filesystem::path& path, const InputGraph& graph);

// Throws std::runtime_error for a missing or unreadable file, wrong magic or
// version, truncation, a boolean byte other than 0/1 or trailing bytes; the
// InputGraph constructor's exceptions for invalid weights or endpoints.
[[nodiscard]] InputGraph read_graph_artifact(const std::filesystem::path& path);

void write_query_artifact(const std::filesystem::path& path, const std::vector<MappedRouteQuery>& queries);

// As read_graph_artifact, and std::runtime_error when a query is not a valid
// RouteQuery of `spatial_graph`, a descriptor count differs from its
// query's terminal count, or a descriptor disagrees with the graph (vertex
// mapping, coordinates, location flag, owner tile) or with its terminal.
[[nodiscard]] std::vector<MappedRouteQuery> read_query_artifact(const std::filesystem::path& path,
                                                                const WeightedGraph& spatial_graph);

// Byte-for-byte equality of two files (std::runtime_error if one cannot be
// opened).
[[nodiscard]] bool files_are_byte_identical(const std::filesystem::path& left,
                                            const std::filesystem::path& right);

}  // namespace bfnew
