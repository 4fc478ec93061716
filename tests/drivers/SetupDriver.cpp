// A generic driver of the host API: it builds and runs the problems that a
// set-up file describes (tests/drivers/setups/*.setup), so that the
// reference's self tests whose gold files are reproduced here are given as
// data -- their parameters, set-up order and the order in which they draw
// from the problem's random number generator -- instead of as programs.
//
// A set-up file is read line by line; '#' starts a comment.  Arguments are
// expressions without blanks: numbers, $names bound by `let`, pi, + - * /
// with the usual precedence (evaluated left to right, as C++ does), unary
// minus, parentheses, and the functions
//   rand(a,b)            the next DPMBase::random.getRandomNumber(a, b)
//   sqrt log sin cos acos square
//   mass(i) radius(i)    of particle i of the handler
//   stiffness()          of the current species
//   collision_time(m) mass_from_radius(r)   of the current species
// Commands:
//   problem dpm|mercury3d|mercury2d|chute   a new problem (ends the previous)
//   name s | dims system particle | gravity x y z | chute_gravity deg g
//   domain x0 y0 z0 x1 y1 z1 | xmax v | ymax v | zmax v
//   timestep e | timemax e | savecount n | hookstride n
//   file data|ene|fstat|restart none|one|multiple
//   seed n | let name e
//   species linear|linear_sliding|linear_friction|hertz_mindlin
//   density | stiffness | dissipation | sliding_stiffness |
//       sliding_dissipation | sliding_friction e       (current species)
//   collision_time_restitution tc e m                  (current species)
//   particle [radius e] [position x y z] [velocity x y z] [fixed]
//   set i [radius e] [position x y z] [velocity x y z] [omega x y z]
//       [orientation q0 q1 q2 q3]       particle i (radius: mass recomputed)
//   wall infinite nx ny nz px py pz
//   boundary deletion nx ny nz d | boundary cube_deletion x0 y0 z0 x1 y1 z1
//   chute length|width|inflow_radius|fixed_radius e | chute bottom TYPE
//   initial ... end     commands run at every setupInitialConditions
//   solve
//   expect_particles n          print the count; exit 1 if it differs
//   expect_velocity i vx vy tol print "ok" if particle i's velocity is within
//                               tol of (vx, vy), else exit 1
#include <cmath>
#include <cstdio>
#include <fstream>
#include <functional>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "Chute.h"

namespace {

struct Context;

// runs the `initial` block of the set-up at every setupInitialConditions,
// after the base class's own set-up
template <class Base>
class Scripted : public Base {
public:
    std::function<void()> initial;
    void setupInitialConditions() override
    {
        Base::setupInitialConditions();
        if (initial) initial();
    }
};

struct Context {
    std::unique_ptr<DPMBase> problem;
    Chute* chute = nullptr;
    std::function<void(std::function<void()>)> setInitial;
    ParticleSpecies* species = nullptr;
    std::map<std::string, double> vars;

    DPMBase& p()
    {
        if (!problem) throw std::runtime_error("no problem defined yet");
        return *problem;
    }
    LinearViscoelasticSpecies& linear()
    {
        auto* s = dynamic_cast<LinearViscoelasticSpecies*>(species);
        if (!s) throw std::runtime_error("the current species is not linear viscoelastic");
        return *s;
    }
    LinearViscoelasticSlidingFrictionSpecies& sliding()
    {
        auto* s = dynamic_cast<LinearViscoelasticSlidingFrictionSpecies*>(species);
        if (!s) throw std::runtime_error("the current species has no sliding friction");
        return *s;
    }
};

// recursive-descent evaluation of one expression
class Expr {
public:
    Expr(const std::string& s, Context& c) : s_(s), c_(c) {}
    double value()
    {
        const double v = sum();
        if (k_ != s_.size()) fail("unexpected '" + s_.substr(k_) + "'");
        return v;
    }

private:
    const std::string& s_;
    Context& c_;
    size_t k_ = 0;

    [[noreturn]] void fail(const std::string& m) { throw std::runtime_error("expression '" + s_ + "': " + m); }
    bool eat(char ch)
    {
        if (k_ < s_.size() && s_[k_] == ch) {
            ++k_;
            return true;
        }
        return false;
    }
    double sum()
    {
        double v = product();
        for (;;) {
            if (eat('+'))
                v = v + product();
            else if (eat('-'))
                v = v - product();
            else
                return v;
        }
    }
    double product()
    {
        double v = unary();
        for (;;) {
            if (eat('*'))
                v = v * unary();
            else if (eat('/'))
                v = v / unary();
            else
                return v;
        }
    }
    double unary()
    {
        if (eat('-')) return -unary();
        if (eat('+')) return unary();
        return primary();
    }
    std::vector<double> args()
    {
        std::vector<double> a;
        if (!eat('(')) fail("'(' expected");
        if (eat(')')) return a;
        do a.push_back(sum());
        while (eat(','));
        if (!eat(')')) fail("')' expected");
        return a;
    }
    double primary()
    {
        if (eat('(')) {
            const double v = sum();
            if (!eat(')')) fail("')' expected");
            return v;
        }
        if (eat('$')) {
            const std::string n = ident();
            auto it = c_.vars.find(n);
            if (it == c_.vars.end()) fail("unknown name $" + n);
            return it->second;
        }
        if (k_ < s_.size() && (std::isdigit((unsigned char)s_[k_]) || s_[k_] == '.')) {
            size_t used = 0;
            const double v = std::stod(s_.substr(k_), &used);
            k_ += used;
            return v;
        }
        const std::string f = ident();
        if (f == "pi") return constants::pi;
        const std::vector<double> a = args();
        auto need = [&](size_t n) {
            if (a.size() != n) fail(f + " takes " + std::to_string(n) + " arguments");
        };
        if (f == "rand") {
            need(2);
            return c_.p().random.getRandomNumber(a[0], a[1]);
        }
        if (f == "sqrt") return need(1), std::sqrt(a[0]);
        if (f == "log") return need(1), std::log(a[0]);
        if (f == "sin") return need(1), std::sin(a[0]);
        if (f == "cos") return need(1), std::cos(a[0]);
        if (f == "acos") return need(1), std::acos(a[0]);
        if (f == "square") return need(1), a[0] * a[0];
        if (f == "mass") return need(1), c_.p().particleHandler.getObject((unsigned)a[0])->getMass();
        if (f == "radius") return need(1), c_.p().particleHandler.getObject((unsigned)a[0])->getRadius();
        if (f == "stiffness") return need(0), c_.linear().getStiffness();
        if (f == "collision_time") return need(1), c_.linear().getCollisionTime(a[0]);
        if (f == "mass_from_radius") {
            need(1);
            if (!c_.species) fail("no species");
            return c_.species->getMassFromRadius(a[0]);
        }
        fail("unknown function " + f);
    }
    std::string ident()
    {
        const size_t b = k_;
        while (k_ < s_.size() && (std::isalnum((unsigned char)s_[k_]) || s_[k_] == '_')) ++k_;
        if (b == k_) fail("name expected");
        return s_.substr(b, k_ - b);
    }
};

FileType fileType(const std::string& s)
{
    if (s == "none") return FileType::NO_FILE;
    if (s == "one") return FileType::ONE_FILE;
    if (s == "multiple") return FileType::MULTIPLE_FILES;
    throw std::runtime_error("unknown file type " + s);
}

RoughBottomType bottomType(const std::string& s)
{
    if (s == "MONOLAYER_ORDERED") return MONOLAYER_ORDERED;
    if (s == "MONOLAYER_DISORDERED") return MONOLAYER_DISORDERED;
    if (s == "MULTILAYER") return MULTILAYER;
    if (s == "FLAT") return FLAT;
    if (s == "MONOLAYER_TRIANGULAR") return MONOLAYER_TRIANGULAR;
    throw std::runtime_error("unknown bottom type " + s);
}

using Line = std::vector<std::string>;

class Runner {
public:
    int run(const std::vector<Line>& lines)
    {
        for (size_t k = 0; k < lines.size(); ++k) {
            const Line& w = lines[k];
            if (w[0] == "initial") {
                std::vector<Line> block;
                while (++k < lines.size() && lines[k][0] != "end") block.push_back(lines[k]);
                if (k == lines.size()) throw std::runtime_error("initial block without end");
                c_.setInitial([this, block]() {
                    for (const Line& b : block) command(b);
                });
                continue;
            }
            const int r = command(w);
            if (r) return r;
        }
        return 0;
    }

private:
    Context c_;

    double num(const std::string& s) { return Expr(s, c_).value(); }
    Vec3D vec(const Line& w, size_t& i)
    {
        if (i + 3 > w.size()) throw std::runtime_error("three values expected after " + w[i - 1]);
        const double x = num(w[i]), y = num(w[i + 1]), z = num(w[i + 2]);
        i += 3;
        return Vec3D(x, y, z);
    }

    template <class B>
    void makeProblem()
    {
        auto* q = new Scripted<B>();
        c_.problem.reset(q);
        c_.chute = dynamic_cast<Chute*>(c_.problem.get());
        c_.setInitial = [q](std::function<void()> f) { q->initial = std::move(f); };
        c_.species = nullptr;
    }

    template <class S>
    void addSpecies()
    {
        c_.species = c_.p().speciesHandler.copyAndAddObject(S());
    }

    // particle options from w[i...] applied to p
    void particleOptions(const Line& w, size_t i, BaseParticle& p, bool recomputeMass)
    {
        while (i < w.size()) {
            const std::string& o = w[i++];
            if (o == "radius") {
                p.setRadius(num(w.at(i++)));
                if (recomputeMass && p.getSpecies()) p.getSpecies()->computeMass(&p);
            } else if (o == "position") {
                p.setPosition(vec(w, i));
            } else if (o == "velocity") {
                p.setVelocity(vec(w, i));
            } else if (o == "omega") {
                p.setAngularVelocity(vec(w, i));
            } else if (o == "orientation") {
                const double q[4] = {num(w.at(i)), num(w.at(i + 1)), num(w.at(i + 2)), num(w.at(i + 3))};
                i += 4;
                p.setOrientation(q);
            } else if (o == "fixed") {
                p.fixParticle();
            } else {
                throw std::runtime_error("unknown particle option " + o);
            }
        }
    }

    int command(const Line& w)
    {
        const std::string& c = w[0];
        auto arg = [&](size_t i) -> const std::string& {
            if (i >= w.size()) throw std::runtime_error("missing argument of " + c);
            return w[i];
        };
        if (c == "problem") {
            const std::string& t = arg(1);
            if (t == "dpm")
                makeProblem<DPMBase>();
            else if (t == "mercury3d")
                makeProblem<Mercury3D>();
            else if (t == "mercury2d")
                makeProblem<Mercury2D>();
            else if (t == "chute")
                makeProblem<Chute>();
            else
                throw std::runtime_error("unknown problem " + t);
        } else if (c == "name") {
            c_.p().setName(arg(1));
        } else if (c == "dims") {
            c_.p().setSystemDimensions((unsigned)num(arg(1)));
            c_.p().setParticleDimensions((unsigned)num(arg(2)));
        } else if (c == "gravity") {
            size_t i = 1;
            c_.p().setGravity(vec(w, i));
        } else if (c == "chute_gravity") {
            if (!c_.chute) throw std::runtime_error("chute_gravity needs a chute problem");
            c_.chute->setChuteAngleAndMagnitudeOfGravity(num(arg(1)), num(arg(2)));
        } else if (c == "domain") {
            size_t i = 1;
            const Vec3D a = vec(w, i), b = vec(w, i);
            c_.p().setDomain(a, b);
        } else if (c == "xmax") {
            c_.p().setXMax(num(arg(1)));
        } else if (c == "ymax") {
            c_.p().setYMax(num(arg(1)));
        } else if (c == "zmax") {
            c_.p().setZMax(num(arg(1)));
        } else if (c == "timestep") {
            c_.p().setTimeStep(num(arg(1)));
        } else if (c == "timemax") {
            c_.p().setTimeMax(num(arg(1)));
        } else if (c == "savecount") {
            c_.p().setSaveCount((unsigned)num(arg(1)));
        } else if (c == "hookstride") {
            c_.p().setHookStride((unsigned)num(arg(1)));
        } else if (c == "file") {
            const std::string& f = arg(1);
            File* file = f == "data" ? &c_.p().dataFile
                         : f == "ene" ? &c_.p().eneFile
                         : f == "fstat" ? &c_.p().fStatFile
                         : f == "restart" ? &c_.p().restartFile
                                          : nullptr;
            if (!file) throw std::runtime_error("unknown file " + f);
            file->setFileType(fileType(arg(2)));
        } else if (c == "seed") {
            c_.p().random.setRandomSeed((unsigned long)num(arg(1)));
        } else if (c == "let") {
            c_.vars[arg(1)] = num(arg(2));
        } else if (c == "species") {
            const std::string& t = arg(1);
            if (t == "linear")
                addSpecies<LinearViscoelasticSpecies>();
            else if (t == "linear_sliding")
                addSpecies<LinearViscoelasticSlidingFrictionSpecies>();
            else if (t == "linear_friction")
                addSpecies<LinearViscoelasticFrictionSpecies>();
            else if (t == "hertz_mindlin")
                addSpecies<HertzianViscoelasticMindlinSpecies>();
            else
                throw std::runtime_error("unknown species " + t);
        } else if (c == "density") {
            if (!c_.species) throw std::runtime_error("no species");
            c_.species->setDensity(num(arg(1)));
        } else if (c == "stiffness") {
            c_.linear().setStiffness(num(arg(1)));
        } else if (c == "dissipation") {
            c_.linear().setDissipation(num(arg(1)));
        } else if (c == "sliding_stiffness") {
            c_.sliding().setSlidingStiffness(num(arg(1)));
        } else if (c == "sliding_dissipation") {
            c_.sliding().setSlidingDissipation(num(arg(1)));
        } else if (c == "sliding_friction") {
            c_.sliding().setSlidingFrictionCoefficient(num(arg(1)));
        } else if (c == "collision_time_restitution") {
            c_.linear().setCollisionTimeAndRestitutionCoefficient(num(arg(1)), num(arg(2)), num(arg(3)));
        } else if (c == "particle") {
            SphericalParticle q(c_.species);
            particleOptions(w, 1, q, false);
            c_.p().particleHandler.copyAndAddObject(q);
        } else if (c == "set") {
            BaseParticle* q = c_.p().particleHandler.getObject((unsigned)num(arg(1)));
            particleOptions(w, 2, *q, true);
        } else if (c == "wall") {
            if (arg(1) != "infinite") throw std::runtime_error("unknown wall " + w[1]);
            size_t i = 2;
            const Vec3D n = vec(w, i), x = vec(w, i);
            InfiniteWall wall;
            wall.setSpecies(c_.p().speciesHandler.getObject(0));
            wall.set(n, x);
            c_.p().wallHandler.copyAndAddObject(wall);
        } else if (c == "boundary") {
            size_t i = 2;
            if (arg(1) == "deletion") {
                const Vec3D n = vec(w, i);
                DeletionBoundary b;
                b.set(n, num(arg(i)));
                c_.p().boundaryHandler.copyAndAddObject(b);
            } else if (arg(1) == "cube_deletion") {
                const Vec3D a = vec(w, i), b = vec(w, i);
                CubeDeletionBoundary d;
                d.set(a, b);
                c_.p().boundaryHandler.copyAndAddObject(d);
            } else {
                throw std::runtime_error("unknown boundary " + w[1]);
            }
        } else if (c == "chute") {
            if (!c_.chute) throw std::runtime_error("chute settings need a chute problem");
            const std::string& s = arg(1);
            if (s == "length")
                c_.chute->setChuteLength(num(arg(2)));
            else if (s == "width")
                c_.chute->setChuteWidth(num(arg(2)));
            else if (s == "inflow_radius")
                c_.chute->setInflowParticleRadius(num(arg(2)));
            else if (s == "fixed_radius")
                c_.chute->setFixedParticleRadius(num(arg(2)));
            else if (s == "bottom")
                c_.chute->setRoughBottomType(bottomType(arg(2)));
            else
                throw std::runtime_error("unknown chute setting " + s);
        } else if (c == "solve") {
            c_.p().solve();
        } else if (c == "expect_particles") {
            const unsigned n = c_.p().particleHandler.getNumberOfObjects();
            std::printf("particles left %u\n", n);
            if (n != (unsigned)num(arg(1))) return 1;
        } else if (c == "expect_velocity") {
            const Vec3D v = c_.p().particleHandler.getObject((unsigned)num(arg(1)))->getVelocity();
            const double tol = num(arg(4));
            const bool ok = std::fabs(v.X - num(arg(2))) < tol && std::fabs(v.Y - num(arg(3))) < tol;
            std::printf("%s: v = (%.9g, %.9g, %.9g) %s\n", c_.p().getName().c_str(), v.X, v.Y, v.Z,
                        ok ? "ok" : "FAILED");
            if (!ok) return 1;
        } else {
            throw std::runtime_error("unknown command " + c);
        }
        return 0;
    }
};

}  // namespace

int main(int argc, char* argv[])
{
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s file.setup\n", argv[0]);
        return 2;
    }
    std::ifstream in(argv[1]);
    if (!in) {
        std::fprintf(stderr, "cannot read %s\n", argv[1]);
        return 2;
    }
    std::vector<Line> lines;
    std::string text;
    while (std::getline(in, text)) {
        const size_t hash = text.find('#');
        if (hash != std::string::npos) text.erase(hash);
        std::istringstream words(text);
        Line w;
        for (std::string s; words >> s;) w.push_back(s);
        if (!w.empty()) lines.push_back(w);
    }
    try {
        return Runner().run(lines);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "%s: %s\n", argv[1], e.what());
        return 2;
    }
}
