// Input decks and the XML basis library (declarations and the reference
// functions each part follows in inputs.hpp).
#include "inputs.hpp"

#include <algorithm>
#include <cctype>
#include <dirent.h>

#include <fstream>
#include <sstream>
#include <stdexcept>

namespace molecular {

namespace {

const char* const kSymbols[] = {
    "h",  "he", "li", "be", "b",  "c",  "n",  "o",  "f",  "ne", "na", "mg", "al", "si", "p",  "s",  "cl", "ar",
    "k",  "ca", "sc", "ti", "v",  "cr", "mn", "fe", "co", "ni", "cu", "zn", "ga", "ge", "as", "se", "br", "kr",
    "rb", "sr", "y",  "zr", "nb", "mo", "tc", "ru", "rh", "pd", "ag", "cd", "in", "sn", "sb", "te", "i",  "xe",
    "cs", "ba", "la", "ce", "pr", "nd", "pm", "sm", "eu", "gd", "tb", "dy", "ho", "er", "tm", "yb", "lu", "hf",
    "ta", "w",  "re", "os", "ir", "pt", "au", "hg", "tl", "pb", "bi", "po", "at", "rn", "fr", "ra", "ac", "th",
    "pa", "u",  "np", "pu", "am", "cm", "bk", "cf", "es", "fm", "md", "no", "lr", "rf", "db", "sg", "bh", "hs",
    "mt"};
constexpr int kNumElements = sizeof(kSymbols) / sizeof(kSymbols[0]);
constexpr double kToBohr = 1.889726124565;   // src/logger.cpp:31

std::string strip_blanks(std::string s) {
    s.erase(std::remove_if(s.begin(), s.end(), [](char c) { return c == ' ' || c == '\t' || c == '\r'; }), s.end());
    return s;
}

DeckOption make_option(const std::string& line) {
    DeckOption o;
    const size_t pos = line.find(',');
    if (pos != std::string::npos) {
        o.name = line.substr(0, pos);
        o.value = line.substr(pos + 1);
        if (o.value == "true" || o.value == "yes" || o.value == "on") o.value = "1";
        else if (o.value == "false" || o.value == "no" || o.value == "off") o.value = "0";
    }
    return o;
}

// Construct::parse: "name={" lines open a sub-construct that runs to its
// matching "}"; everything else is content
Construct make_construct(const std::vector<std::string>& lines, const std::string& name) {
    Construct c;
    c.name = name;
    for (size_t i = 0; i < lines.size(); ++i) {
        size_t pos = lines[i].find('=');
        if (pos == std::string::npos) {
            c.content.push_back(lines[i]);
            continue;
        }
        const std::string token = lines[i].substr(0, pos);
        std::vector<std::string> sub;
        ++i;
        bool end_found = i >= lines.size() || lines[i].find('}') != std::string::npos;
        int nsubs = 0;
        while (i < lines.size() && !end_found) {
            sub.push_back(lines[i]);
            if (lines[i].find('=') != std::string::npos) {
                ++nsubs;
                ++i;
            } else {
                const bool close = lines[i++].find('}') != std::string::npos;
                if (close && nsubs == 0) {
                    end_found = true;
                    sub.pop_back();
                    --i;
                } else if (close) {
                    --nsubs;
                }
            }
        }
        c.sub.push_back(make_construct(sub, token));
    }
    return c;
}

// Basis names of one basis construct (Molecule::parseBasis)
void parse_basis_names(const Construct& sc, std::map<int, std::string>& names) {
    for (const std::string& line : sc.content) {
        const size_t pos = line.find(',');
        if (pos == std::string::npos) continue;
        const std::string token = line.substr(0, pos), name = line.substr(pos + 1);
        names[token == "default" ? 0 : element_charge(token)] = name;
    }
}

// Molecule::parseGeomLine: "symbol, x, y, z"
Atom parse_geom_line(const std::string& line, bool angstrom) {
    std::vector<std::string> tok;
    std::stringstream ss(line);
    std::string t;
    while (std::getline(ss, t, ',')) tok.push_back(strip_blanks(t));
    if (tok.size() < 4 || tok[3].empty()) throw std::runtime_error("INPUT: Missing coordinate in '" + line + "'");
    const double m = angstrom ? kToBohr : 1.0;
    Atom a;
    a.Z = element_charge(tok[0]);
    a.x = m * std::stod(tok[1]);
    a.y = m * std::stod(tok[2]);
    a.z = m * std::stod(tok[3]);
    return a;
}

std::string read_file(const std::string& path) {
    std::ifstream in(path);
    if (!in) return std::string();
    std::stringstream ss;
    ss << in.rdbuf();
    return ss.str();
}

// value of attribute `name` inside the tag text [b, e) of an XML document
bool xml_attr(const std::string& doc, size_t b, size_t e, const std::string& name, std::string& out) {
    const std::string key = name + "=\"";
    size_t p = b;
    while ((p = doc.find(key, p)) != std::string::npos && p < e) {
        if (p > b && (std::isspace((unsigned char)doc[p - 1]) != 0)) {
            const size_t v = p + key.size();
            const size_t q = doc.find('"', v);
            if (q == std::string::npos || q > e) return false;
            out = doc.substr(v, q - v);
            return true;
        }
        p += key.size();
    }
    return false;
}

// start of the opening tag <tag ...> or <tag> at or after `from` (whole tag name only)
size_t xml_open(const std::string& doc, const std::string& tag, size_t from, size_t until) {
    const std::string key = "<" + tag;
    size_t p = from;
    while ((p = doc.find(key, p)) != std::string::npos && p < until) {
        const size_t n = p + key.size();
        if (n < doc.size() && (doc[n] == '>' || doc[n] == '/' || std::isspace((unsigned char)doc[n]) != 0)) return p;
        p = n;
    }
    return std::string::npos;
}

int lval(const std::string& s) {
    // BasisReader::read_basis (src/basisreader.cpp:62-74), "sp" read as l = 1 as there
    static const std::map<std::string, int> m = {{"s", 0}, {"sp", 1}, {"p", 1}, {"d", 2}, {"f", 3}, {"g", 4},
                                                  {"h", 5}, {"i", 6}, {"k", 7}, {"l", 8}};
    auto it = m.find(s);
    return it == m.end() ? 0 : it->second;
}

}  // namespace

const std::string* DeckCommand::find(const std::string& key) const {
    for (const DeckOption& o : options)
        if (o.name == key) return &o.value;
    return nullptr;
}

const std::string* Deck::global(const std::string& key) const {
    for (const DeckOption& o : globals)
        if (o.name == key) return &o.value;
    return nullptr;
}

void clean_line(std::string& line) {
    const size_t pos = line.find('!');
    if (pos != std::string::npos) line.erase(pos);
    line = strip_blanks(line);
    std::transform(line.begin(), line.end(), line.begin(), [](unsigned char c) { return (char)std::tolower(c); });
}

Deck parse_deck(const std::string& text) {
    Deck deck;
    std::istringstream in(text);
    std::string line;
    int curr_molecule = -1;
    while (std::getline(in, line)) {
        clean_line(line);
        size_t pos = line.find('=');
        if (pos != std::string::npos) {   // construct
            const std::string token = line.substr(0, pos);
            std::vector<std::string> sub;
            if (std::getline(in, line)) {
                bool end_found = line.find('}') != std::string::npos;
                int nsubs = 0;
                while (!end_found) {
                    clean_line(line);
                    sub.push_back(line);
                    if (line.find('=') != std::string::npos) {
                        ++nsubs;
                        if (!std::getline(in, line)) end_found = true;
                    } else {
                        const bool close = line.find('}') != std::string::npos;
                        if (close && nsubs == 0) {
                            end_found = true;
                            sub.pop_back();
                        } else if (close) {
                            --nsubs;
                            if (!std::getline(in, line)) end_found = true;
                        } else if (!std::getline(in, line)) {
                            end_found = true;
                        }
                    }
                }
            }
            deck.constructs.push_back(make_construct(sub, token));
            ++curr_molecule;
        } else if ((pos = line.find('(')) != std::string::npos) {   // command
            DeckCommand cmd;
            cmd.name = line.substr(0, pos);
            cmd.molecule = curr_molecule;
            if (line.find(')') == std::string::npos && std::getline(in, line)) {
                while (true) {
                    if (line.find(')') != std::string::npos) break;
                    clean_line(line);
                    cmd.options.push_back(make_option(line));
                    if (!std::getline(in, line)) break;
                }
            }
            deck.commands.push_back(cmd);
        } else if (line.find(',') != std::string::npos) {   // global option
            deck.globals.push_back(make_option(line));
        }
    }
    return deck;
}

int element_charge(const std::string& symbol) {
    std::string s = strip_blanks(symbol);
    std::transform(s.begin(), s.end(), s.begin(), [](unsigned char c) { return (char)std::tolower(c); });
    for (int q = 0; q < kNumElements; ++q)
        if (s == kSymbols[q]) return q + 1;
    throw std::runtime_error("INPUT: unknown element '" + symbol + "'");
}

std::string element_symbol(int Z) {
    if (Z < 1 || Z > kNumElements) throw std::runtime_error("INPUT: no element with Z = " + std::to_string(Z));
    return kSymbols[Z - 1];
}

DeckMolecule deck_molecule(const Construct& c) {
    DeckMolecule m;
    for (const std::string& line : c.content) {
        const DeckOption o = make_option(line);
        if (o.name == "charge") m.charge = std::stoi(o.value);
        else if (o.name == "multiplicity") m.multiplicity = std::stoi(o.value);
        else if (o.name == "angstrom") m.angstrom = true;   // any value, as Molecule::init
    }
    bool geom = false, basis = false;
    for (const Construct& sc : c.sub) {
        if (sc.name == "basis" || sc.name == "jkbasis" || sc.name == "ribasis") {
            if (sc.content.empty()) throw std::runtime_error("INIT: Basis specification is empty!");
            parse_basis_names(sc, sc.name == "basis" ? m.basis : sc.name == "jkbasis" ? m.jkbasis : m.ribasis);
            basis = basis || sc.name == "basis";
        } else if (sc.name == "geometry") {
            if (sc.content.empty()) throw std::runtime_error("INIT: There are no atoms!");
            for (const std::string& line : sc.content) m.atoms.push_back(parse_geom_line(line, m.angstrom));
            geom = true;
        }
        // fragments, ecp, frozen: read by the reference's ALMO / ECP paths only
    }
    if (!geom) throw std::runtime_error("INIT: molecule has no geometry");
    if (!basis) throw std::runtime_error("INIT: molecule has no basis");
    return m;
}

std::map<std::string, int> read_basis_list(const std::string& libdir, const std::string& kind) {
    std::map<std::string, int> out;
    const std::string doc = read_file(libdir + "/" + kind + ".list");
    size_t p = 0;
    while ((p = xml_open(doc, "Item", p, doc.size())) != std::string::npos) {
        const size_t e = doc.find('>', p);
        if (e == std::string::npos) break;
        std::string key, value;
        if (xml_attr(doc, p, e, "key", key) && xml_attr(doc, p, e, "value", value)) out[key] = std::stoi(value);
        p = e;
    }
    return out;
}

bool is_molpro_name(const std::string& name) {
    return name.size() > 6 && name.compare(name.size() - 6, 6, ".basis") == 0;
}

std::string resolve_basis_file(const std::string& libdir, const std::string& name) {
    // as given, beside the XML library, or in the basissets/ directory that
    // sits beside basislibrary/ in the reference's tree.  Deck text is read
    // in lower case (as the reference parses it), so a directory entry whose
    // name matches apart from case is taken as well.
    auto lower = [](std::string x) {
        std::transform(x.begin(), x.end(), x.begin(), [](unsigned char c) { return (char)std::tolower(c); });
        return x;
    };
    for (const std::string& p : {name, libdir + "/" + name, libdir + "/../basissets/" + name}) {
        std::ifstream f(p);
        if (f.good()) return p;
    }
    const size_t slash = name.find_last_of('/');
    const std::string base = lower(slash == std::string::npos ? name : name.substr(slash + 1));
    const std::string given = slash == std::string::npos ? std::string(".") : name.substr(0, slash);
    for (const std::string& dir : {given, libdir, libdir + "/../basissets"}) {
        DIR* d = opendir(dir.c_str());
        if (!d) continue;
        std::string hit;
        while (dirent* e = readdir(d))
            if (lower(e->d_name) == base) {
                hit = dir + "/" + e->d_name;
                break;
            }
        closedir(d);
        if (!hit.empty()) return hit;
    }
    throw std::runtime_error("BASISLIB: Could not open basis file " + name);
}

// Molpro basis text: "<l>, <element>, e1, e2, ..." gives the exponents of an
// angular momentum; each following "c, i.j, c_i, ..., c_j" line is one
// segmented shell over exponents i..j (1-based), which is how the
// reference's basis library splits general contractions into shells
// (src/basisreader.cpp:62-86 reads the XML built from this text with one
// <Shell> per contraction line).  Lines starting with '!' are comments.
std::vector<RawShell> molpro_element_shells(const std::string& text, const std::string& symbol) {
    auto lower = [](std::string x) {
        std::transform(x.begin(), x.end(), x.begin(), [](unsigned char c) { return (char)std::tolower(c); });
        return x;
    };
    const std::string want = lower(symbol);
    std::vector<RawShell> out;
    std::vector<double> exps;
    int l = -1;
    bool mine = false;
    std::istringstream in(text);
    std::string line;
    while (std::getline(in, line)) {
        const size_t b = line.find_first_not_of(" \t\r");
        if (b == std::string::npos || line[b] == '!' || line[b] == '}' || lower(line.substr(b, 5)) == "basis")
            continue;
        std::vector<std::string> f;
        std::stringstream ss(line.substr(b));
        std::string item;
        while (std::getline(ss, item, ',')) f.push_back(strip_blanks(item));
        if (f.size() < 2) continue;
        const std::string head = lower(f[0]);
        if (head == "c") {
            if (!mine) continue;
            const size_t dot = f[1].find('.');
            if (l < 0 || dot == std::string::npos) throw std::runtime_error("BASISLIB: malformed Molpro line: " + line);
            const int lo = std::stoi(f[1].substr(0, dot)), hi = std::stoi(f[1].substr(dot + 1));
            if (lo < 1 || hi > (int)exps.size() || hi < lo || (int)f.size() - 2 != hi - lo + 1)
                throw std::runtime_error("BASISLIB: contraction does not fit its exponents: " + line);
            RawShell s;
            s.l = l;
            for (int k = lo; k <= hi; ++k) {
                s.exps.push_back(exps[k - 1]);
                s.coefs.push_back(std::stod(f[2 + k - lo]));
            }
            out.push_back(s);
        } else if (head.size() == 1 && std::string("spdfghi").find(head[0]) != std::string::npos) {
            l = (int)std::string("spdfghi").find(head[0]);
            mine = lower(f[1]) == want;
            exps.clear();
            if (mine)
                for (size_t k = 2; k < f.size(); ++k)
                    if (!f[k].empty()) exps.push_back(std::stod(f[k]));
        }
    }
    return out;
}

std::vector<RawShell> library_shells(const std::string& libdir, const std::string& kind,
                                     const std::vector<Atom>& atoms, const std::map<int, std::string>& names) {
    const std::map<std::string, int> list = read_basis_list(libdir, kind);
    std::map<int, std::string> docs;   // file number -> text
    std::map<std::string, std::string> molpro;   // Molpro file name -> text
    std::vector<RawShell> out;
    for (size_t ia = 0; ia < atoms.size(); ++ia) {
        const Atom& a = atoms[ia];
        std::string name = "sto-3g";
        auto it = names.find(a.Z);
        if (it != names.end()) name = it->second;
        else if ((it = names.find(0)) != names.end()) name = it->second;
        const std::string sym = element_symbol(a.Z);
        if (is_molpro_name(name)) {
            // a Molpro basis file named directly in the deck
            if (!molpro.count(name)) molpro[name] = read_file(resolve_basis_file(libdir, name));
            std::vector<RawShell> sh = molpro_element_shells(molpro[name], sym);
            if (sh.empty()) throw std::runtime_error("BASISLIB: Could not read basis " + name + " for atom " + sym);
            for (RawShell& s : sh) {
                s.atom = (int)ia;
                s.O[0] = a.x;
                s.O[1] = a.y;
                s.O[2] = a.z;
                out.push_back(s);
            }
            continue;
        }
        std::transform(name.begin(), name.end(), name.begin(), [](unsigned char c) { return (char)std::tolower(c); });
        size_t before = out.size();
        auto li = list.find(name);
        if (li != list.end()) {
            if (!docs.count(li->second))
                docs[li->second] = read_file(libdir + "/bases/build/" + kind + std::to_string(li->second) + ".xml");
            const std::string& doc = docs[li->second];
            const size_t root = xml_open(doc, "root", 0, doc.size());
            const size_t el = root == std::string::npos ? root : xml_open(doc, sym, root + 5, doc.size());
            if (el != std::string::npos) {
                size_t end = doc.find("</" + sym + ">", el);
                if (end == std::string::npos) end = doc.size();
                size_t sp = el;
                while ((sp = xml_open(doc, "Shell", sp + 1, end)) != std::string::npos) {
                    const size_t tag_end = doc.find('>', sp);
                    size_t shell_end = doc.find("</Shell>", sp);
                    if (tag_end == std::string::npos) break;
                    if (shell_end == std::string::npos || shell_end > end) shell_end = end;
                    std::string lv;
                    xml_attr(doc, sp, tag_end, "lval", lv);
                    RawShell s;
                    s.l = lval(lv);
                    s.atom = (int)ia;
                    s.O[0] = a.x;
                    s.O[1] = a.y;
                    s.O[2] = a.z;
                    size_t xp = tag_end;
                    while ((xp = xml_open(doc, "xc", xp, shell_end)) != std::string::npos) {
                        const size_t xe = doc.find('>', xp);
                        std::string x, c;
                        if (xe == std::string::npos || !xml_attr(doc, xp, xe, "x", x) || !xml_attr(doc, xp, xe, "c", c))
                            throw std::runtime_error("BASISLIB: malformed <xc> in " + kind + " basis " + name);
                        s.exps.push_back(std::stod(x));
                        s.coefs.push_back(std::stod(c));
                        xp = xe;
                    }
                    if (!s.exps.empty()) out.push_back(s);
                    sp = shell_end;
                }
            }
        }
        if (out.size() == before)
            throw std::runtime_error("BASISLIB: Could not read basis " + name + " for atom " + sym);
    }
    return out;
}

}  // namespace molecular
