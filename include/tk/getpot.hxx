/*
 * tk/getpot.hxx -- the part of GetPot that rsfe_splitbregman's parameter files use.
 *
 * The reference reads its parameters with GetPot (http://getpot.sourceforge.net/, a
 * third-party header the reference does not ship and pins no version of:
 * rsfe_splitbregman.hxx:28, ReadMe.md:31-38), through
 * GetPot ifile(file, "#", "\n"); ifile.set_prefix("splitbregman/" + section);
 * ifile(name, default) (rsfe_splitbregman_imp.hxx:582-796).  This restates GetPot 1.x's
 * published input-file rules for what those calls need:
 *  - comments run from the comment start ("#") to the comment end ("\n");
 *  - tokens are separated by white space; 'single quotes' keep one token with spaces;
 *  - "name = value", "name= value", "name =value" and "name=value" define name;
 *  - [label] sets the section: an absolute label, "./sub" appends sub, each leading "../"
 *    goes up one level, [] is the root; section paths end with '/' and prefix the names
 *    defined under them;
 *  - a later definition of a name replaces an earlier one;
 *  - ifile(name, default) looks up prefix + name: an int or double default asks for a
 *    number (sscanf "%lf"; ints through double, so 2.0e1 reads 20), and the default comes
 *    back when the name is missing or not a number; a string default returns the text.
 * A missing file gives no variables (every lookup returns its default), as in GetPot.
 */
#ifndef TK_GETPOT_HXX
#define TK_GETPOT_HXX

#include <cctype>
#include <cstdio>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

namespace tk {

class getpot {
public:
    getpot() = default;

    explicit getpot(const char* file, const char* comment_start = "#",
                    const char* comment_end = "\n")
    {
        std::ifstream in(file, std::ios::binary);
        if (!in) return;
        std::stringstream ss;
        ss << in.rdbuf();
        parse(strip_comments(ss.str(), comment_start, comment_end));
    }

    // from text (same rules as a file)
    static getpot from_string(const std::string& text, const char* comment_start = "#",
                              const char* comment_end = "\n")
    {
        getpot g;
        g.parse(strip_comments(text, comment_start, comment_end));
        return g;
    }

    void set_prefix(const char* prefix) { prefix_ = prefix ? prefix : ""; }

    int operator()(const char* name, int dflt) const
    {
        double v;
        return number(name, &v) ? (int)v : dflt;
    }
    double operator()(const char* name, double dflt) const
    {
        double v;
        return number(name, &v) ? v : dflt;
    }
    std::string operator()(const char* name, const char* dflt) const
    {
        auto it = vars_.find(prefix_ + name);
        return it == vars_.end() ? std::string(dflt) : it->second;
    }

    bool have(const char* name) const { return vars_.count(prefix_ + name) != 0; }
    std::map<std::string, std::string> const& variables() const { return vars_; }

private:
    static std::string strip_comments(const std::string& s, const char* cs, const char* ce)
    {
        const std::string a(cs), b(ce);
        std::string out;
        size_t p = 0;
        while (p < s.size()) {
            const size_t q = a.empty() ? std::string::npos : s.find(a, p);
            if (q == std::string::npos) {
                out.append(s, p, std::string::npos);
                break;
            }
            out.append(s, p, q - p);
            const size_t e = s.find(b, q + a.size());
            if (e == std::string::npos) break;
            out.push_back(' ');
            p = e + b.size();
        }
        return out;
    }

    static std::vector<std::string> tokens(const std::string& s)
    {
        std::vector<std::string> t;
        size_t p = 0;
        while (p < s.size()) {
            while (p < s.size() && std::isspace((unsigned char)s[p])) ++p;
            if (p >= s.size()) break;
            std::string tok;
            while (p < s.size() && !std::isspace((unsigned char)s[p])) {
                if (s[p] == '\'') {
                    const size_t e = s.find('\'', p + 1);
                    const size_t stop = e == std::string::npos ? s.size() : e;
                    tok.append(s, p + 1, stop - p - 1);
                    p = stop + (e == std::string::npos ? 0 : 1);
                } else {
                    tok.push_back(s[p++]);
                }
            }
            t.push_back(tok);
        }
        return t;
    }

    void enter_section(const std::string& label)
    {
        if (label.empty()) {
            section_.clear();
            return;
        }
        std::string rest = label;
        if (rest.compare(0, 2, "./") == 0) {
            rest = rest.substr(2);
        } else if (rest.compare(0, 3, "../") == 0) {
            while (rest.compare(0, 3, "../") == 0) {
                rest = rest.substr(3);
                if (!section_.empty()) {
                    const size_t cut = section_.find_last_of('/', section_.size() - 2);
                    section_ = cut == std::string::npos ? "" : section_.substr(0, cut + 1);
                }
            }
        } else {
            section_.clear();
        }
        if (!rest.empty()) {
            section_ += rest;
            if (section_.back() != '/') section_ += '/';
        }
    }

    void parse(const std::string& text)
    {
        const std::vector<std::string> t = tokens(text);
        for (size_t i = 0; i < t.size(); ++i) {
            const std::string& w = t[i];
            if (w.size() >= 2 && w.front() == '[' && w.back() == ']') {
                enter_section(w.substr(1, w.size() - 2));
                continue;
            }
            // join "name = value" / "name= value" / "name =value" into name=value
            std::string name, value;
            const size_t eq = w.find('=');
            if (eq != std::string::npos && eq > 0) {
                name = w.substr(0, eq);
                value = w.substr(eq + 1);
                if (value.empty() && i + 1 < t.size()) value = t[++i];
            } else if (i + 1 < t.size() && !t[i + 1].empty() && t[i + 1][0] == '=') {
                name = w;
                value = t[i + 1].substr(1);
                ++i;
                if (value.empty() && i + 1 < t.size()) value = t[++i];
            } else {
                continue;  // a lone argument: not a variable
            }
            vars_[section_ + name] = value;
        }
    }

    bool number(const char* name, double* v) const
    {
        auto it = vars_.find(prefix_ + name);
        if (it == vars_.end()) return false;
        return std::sscanf(it->second.c_str(), "%lf", v) == 1;
    }

    std::map<std::string, std::string> vars_;
    std::string section_, prefix_;
};

}  // namespace tk

#endif
