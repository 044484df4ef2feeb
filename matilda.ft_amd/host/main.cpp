// Command-line driver of the drop-in front end (main.cu:40-151): reads the
// input file line by line; `box` creates a box and hands it the following
// lines up to endBox; `run nvt|vt N` runs it; `modify` edits it.
//
//   mft [-in input] [-device k] [--dry-run] [--dump-topology file]
//
// --dry-run parses the input and prints the engine description as JSON
// instead of creating the engine (no GPU needed).  --dump-topology writes the
// particle box's topology ("species charge molecule nBonds bondedTo[MAXBONDS]
// bondType[MAXBONDS]" per particle) to file.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <iostream>

#include "matilda.h"

bool Box::dry_run = false;
std::string Box::topology_dump;

void die(const std::string& msg)
{
    std::cout << msg << std::endl;
    std::fflush(stdout);
    std::exit(1);
}

void check_mft(int rc, const std::string& what)
{
    if (rc != MFT_OK) die(what + " failed: " + mft_last_error());
}

// xorshift64* (Vigna, "An experimental exploration of Marsaglia's xorshift
// generators, scrambled", 2016); the top 53 bits give a double in (0, 1)
double HostRng::uniform()
{
    s_ ^= s_ >> 12;
    s_ ^= s_ << 25;
    s_ ^= s_ >> 27;
    uint64_t r = s_ * 0x2545F4914F6CDD1Dull;
    return ((double)(r >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

// uniform direction: z uniform in [-1, 1], azimuth uniform (Archimedes)
void HostRng::unit_vector(double* r)
{
    const double z = 2.0 * uniform() - 1.0;
    const double phi = 6.283185307179586 * uniform();
    const double s = std::sqrt(std::max(0.0, 1.0 - z * z));
    r[0] = s * std::cos(phi);
    r[1] = s * std::sin(phi);
    r[2] = z;
}

void Box::read_grid(std::istringstream& iss)
{
    if (!readDimension) die("Dim must be defined before grid!");
    for (int j = 0; j < Dim; ++j) {
        if (!(iss >> Nx[j])) die("Failed to read grid value! Did you include values for each dimension?");
        M *= Nx[j];
    }
}

void Box::read_lengths(std::istringstream& iss)
{
    if (!readDimension)
        std::cout << "\nWARNING! WARNING!\nboxLengths read before dimension specified, assuming default value of 2"
                  << std::endl;
    for (int j = 0; j < Dim; ++j) {
        if (!(iss >> L[j])) die("Failed to read boxlength value! Did you include values for each dimension?");
        V *= L[j];
    }
}

void Box::findSpinodal(std::istringstream& iss)
{
    (void)iss;
    die("findSpinodal is not part of this build");
}

void Box::modifyBox(std::istringstream& iss)
{
    (void)iss;
    die("modify is not part of this build");
}

Box* BoxFactory(std::istringstream& iss)
{
    std::string s1;
    iss >> s1;
    if (s1 == "fts") return new FTS_Box(iss);
    if (s1 == "ps") return new PS_Box(iss);
    die(s1 + " is not a valid box style yet");
}

int main(int argc, char** argv)
{
    std::string input_file = "input";
    for (int a = 1; a < argc; ++a) {
        std::string s = argv[a];
        if (s == "-in" && a + 1 < argc) {
            input_file = argv[++a];
            std::cout << "Operating from input file " << input_file << std::endl;
        } else if (s == "-device" && a + 1 < argc) {
            int dev = std::atoi(argv[++a]);
            int n = 0;
            check_mft(mft_device_count(&n), "mft_device_count");
            if (dev != 0 || n < 1) die("Device " + std::to_string(dev) + " not found; this build drives device 0");
        } else if (s == "--dry-run") {
            Box::dry_run = true;
        } else if (s == "--dump-topology" && a + 1 < argc) {
            Box::topology_dump = argv[++a];
        } else {
            die(s + " is not a valid argument");
        }
    }
    std::ifstream in(input_file);
    if (!in) die("Failed to open " + input_file);
    std::vector<Box*> box;
    std::string line, word;
    while (std::getline(in, line)) {
        if (line.empty() || line[0] == '#') continue;
        std::istringstream iss(line);
        if (!(iss >> word)) continue;
        if (word == "box") {
            box.push_back(BoxFactory(iss));
            box.back()->readInput(in);
        } else if (word == "run") {
            if (box.empty()) die("run before any box");
            iss >> word;
            int steps = 0;
            if (word == "nvt") {
                iss >> steps;
                box[0]->NVT(steps);
            } else if (word == "vt") {
                if (box[0]->returnBoxStyle() != "fts") die("invalid box style trying to run const 'vt' simulation");
                iss >> steps;
                box[0]->NVT(steps);
            } else if (word == "findspinodal" || word == "findSpinodal") {
                if (box[0]->returnBoxStyle() != "fts") die("invalid box style trying to run const 'vt' simulation");
                box[0]->findSpinodal(iss);
            } else {
                die(word + " is an invalid global command!");
            }
        } else if (word == "modify") {
            if (box.empty()) die("modify before any box");
            box[0]->modifyBox(iss);
        }
    }
    for (Box* b : box) delete b;
    return 0;
}
