// The reference's input files, read on the host: `.mol` input decks
// (ProgramController::parse / Construct::parse, src/ProgramController.cpp:
// 42-236; Molecule::init / parseGeomLine / parseBasis, src/molecule.cpp:
// 20-240) and its XML basis library (BasisReader, src/basisreader.cpp:27-132:
// basislibrary/<kind>.list maps a basis name to a file number N, and
// basislibrary/bases/build/<kind>N.xml holds one element per atom with its
// <Shell lval=".."><xc x=".." c=".."/>...</Shell> list).
#pragma once
#include <map>
#include <string>
#include <vector>

#include "scf_host.hpp"

namespace molecular {

// "name, value" (Option::parse): true/yes/on -> "1", false/no/off -> "0"
struct DeckOption {
    std::string name, value;
};

// "name = { ... }" with nested sub-constructs (Construct::parse)
struct Construct {
    std::string name;
    std::vector<std::string> content;
    std::vector<Construct> sub;
};

// "name( options )"; molecule = index of the last construct before it (-1: none)
struct DeckCommand {
    std::string name;
    int molecule = -1;
    std::vector<DeckOption> options;
    const std::string* find(const std::string& key) const;
};

struct Deck {
    std::vector<DeckOption> globals;
    std::vector<DeckCommand> commands;
    std::vector<Construct> constructs;
    const std::string* global(const std::string& key) const;
};

// ProgramController::cleanLine: drop "!" comments, spaces and tabs; lower case
void clean_line(std::string& line);
// ProgramController::parse over the whole text of a deck; unknown command
// names are kept (the caller decides what runs), as the reference logs them
Deck parse_deck(const std::string& text);

// element symbol <-> nuclear charge (H..Mt); throws for an unknown symbol
int element_charge(const std::string& symbol);
std::string element_symbol(int Z);

struct DeckMolecule {
    int charge = 0, multiplicity = 1;
    bool angstrom = false;
    std::vector<Atom> atoms;                           // bohr
    std::map<int, std::string> basis, jkbasis, ribasis;   // Z -> name, 0 = default
};
// Molecule::init over a "molecule" construct (geometry, basis, jkbasis,
// ribasis; charge, multiplicity, angstrom); throws on a malformed deck
DeckMolecule deck_molecule(const Construct& c);

// BasisReader::read_basis_list: basis name -> library file number
std::map<std::string, int> read_basis_list(const std::string& libdir, const std::string& kind);
// BasisReader::readShellBasis for every atom, in atom order: the basis named
// for the atom's element (else the default, else sto-3g) from
// <libdir>/bases/build/<kind>N.xml; throws BASISLIB when an atom has no shells
std::vector<RawShell> library_shells(const std::string& libdir, const std::string& kind,
                                     const std::vector<Atom>& atoms, const std::map<int, std::string>& names);
// A basis name ending in ".basis" names a Molpro basis text file instead of
// a library entry; it is found as given, in <libdir>, or in
// <libdir>/../basissets (the reference's basissets/ directory).
bool is_molpro_name(const std::string& name);
std::string resolve_basis_file(const std::string& libdir, const std::string& name);
// the shells of one element from Molpro basis text, in file order, one
// segmented shell per contraction line (no atom or centre set)
std::vector<RawShell> molpro_element_shells(const std::string& text, const std::string& symbol);

}  // namespace molecular
