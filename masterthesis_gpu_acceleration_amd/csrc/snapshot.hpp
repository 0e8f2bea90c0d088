// snapshot.hpp -- binary snapshot of a built graph: the predicate-grouped,
// sorted adjacency arrays plus the term dictionary and prefix lines, so a
// graph is loaded again without parsing Turtle or sorting triples.
#pragma once
#include <string>

#include "graph_host.hpp"

namespace gr {

// Writes g to `path`. False (with err) if the file cannot be written.
bool save_snapshot(const HostGraph& g, const std::string& path, std::string& err);

// Reads a snapshot written by save_snapshot into g. The file is mapped and
// checked before anything is kept: section sizes against the file, term
// ids against the row space, per-predicate sort order of both directions,
// and that the POS arrays hold the same triples as the SPO arrays. False
// (with err) on any mismatch.
bool load_snapshot(const std::string& path, HostGraph& g, std::string& err);

}  // namespace gr
