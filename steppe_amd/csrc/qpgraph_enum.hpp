// steppe_amd/csrc/qpgraph_enum.hpp — the bounded qpGraph topology space: every rooted
// binary tree on a leaf set and every graph with one admixture event, the isomorphism
// key that de-duplicates them, and the local moves of the hill-climb (host code;
// qpgraph_enum.cpp).
//
// Behaviour of the reference's enumerator (src/core/qpadm/qpgraph_enumerate.cpp):
// candidate order, node names ("N<id>"), edge order within a graph and the 64-bit
// shape hash are the reference's, so candidate ids and hashes are interchangeable with
// its results and with the admixtools scores keyed by them.
#pragma once

#include <cstdint>
#include <string>
#include <utility>
#include <vector>

namespace stp::qgenum {

using NamedEdges = std::vector<std::pair<std::string, std::string>>;

struct Candidate {
    NamedEdges edges;
    int nadmix = 0;
    int id = 0;
    std::uint64_t hash = 0;
};

// Trees (nadmix 0) in insertion order, then (max_nadmix >= 1) the distinct
// one-admixture graphs wired onto each tree in turn (qpgraph_enumerate.cpp:205-236).
std::vector<Candidate> enumerate_space(const std::vector<std::string>& leaves, int max_nadmix);

// The shape key of a labelled edge list: leaves (nodes without children) ranked by
// label (qpgraph_enumerate.cpp:183-202).
std::uint64_t graph_hash(const NamedEdges& edges);

// Hill-climb moves from a candidate: NNI trees of its base tree, then (max_nadmix >= 1)
// admixture wirings of the base tree and of each NNI tree (qpgraph_enumerate.cpp:383-407).
std::vector<Candidate> neighbors(const Candidate& current, const std::vector<std::string>& leaves,
                                 int max_nadmix);

}  // namespace stp::qgenum
