// inter.dat: the hand-off between `process` and `simulate`.
//
// The reference writes a Boost.Serialization text archive of its `inter`
// object (src/process.cpp:23-81, written at :604-606, read at
// src/simulate.cu:1835-1838).  This build writes the same object, in the
// same member order and text-archive syntax, derived from the parsed design
// (readers that stop after `inter` ignore what follows), and appends a
// section of its own: the cell library, the netlist and the levelized,
// compiled program, so `simulate` starts from the compiled program instead
// of parsing and levelizing again.
#pragma once
#include <memory>
#include <string>

#include "csr.h"
#include "netlist.h"

namespace gls {

// Text of the `inter` archive (the reference's member order) for the design
// and its program compiled against the netlist's declared inputs.
std::string inter_archive_text(const Library& lib, const Netlist& nl, const Program& pg);

// Source texts of a parsed design (all empty for a generated one).
struct DesignTexts {
  std::string vlib, gv, sdf;
};

// Whole inter.dat: archive text, then the compiled section (library,
// netlist, program, source texts).
std::string inter_file_text(const Library& lib, const Netlist& nl, const Program& pg,
                            const DesignTexts& texts);

// Reads the compiled section of an inter.dat written by inter_file_text.
// Returns false when the text has no such section.  Throws
// GlsError(GLS_E_FORMAT) for a damaged section.
bool read_inter_section(const std::string& text, Library& lib, Netlist& nl, Program& pg,
                        DesignTexts& texts);

}  // namespace gls
