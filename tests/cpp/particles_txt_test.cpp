// Text restart format of wacomm_hip::Particles (Particles.cpp:28-55, 320-364),
// restating tests/RestartTest.cpp:104-122 of WaComM++ plus the reader's
// other branches.  CPU only: no device call is made.
#include <cassert>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>

#include "wacomm_hip.hpp"

using wacomm_hip::Particle;
using wacomm_hip::Particles;

int main() {
    const std::string textFile = "wc-restart-test.txt";
    // tests/RestartTest.cpp:104-122: exact 64-bit identity and full double precision
    {
        Particles written;
        Particle particle(9007199254740993ULL, -.123456789012345, 5.85472728621506, 5.32375564640966,
                          .920044414487148, 7200.0000133514404, 1675900800.25);
        particle.Drift(1, 1);                      // PERSON_IN_WATER, RIGHT (DriftModel.hpp:14-44)
        written.push_back(particle);
        written.saveAsTxt(textFile);
        Particles read;
        read.loadFromTxt(textFile);
        assert(read.size() == 1);
        assert(read[0].Id() == 9007199254740993ULL);
        assert(read[0].K() == particle.K() && read[0].J() == particle.J() && read[0].I() == particle.I());
        assert(read[0].Health() == particle.Health() && read[0].Age() == particle.Age());
        assert(read[0].Time() == particle.Time());
        assert(read[0].DriftObject() == 1 && read[0].Side() == 1);
    }
    // round trip of awkward doubles: every value must come back bit for bit
    {
        Particles written;
        const double vals[] = {1.0 / 3.0, -0.0, 5e-324, 1.7976931348623157e308, -2.5e-310, 0.1, 1e16 + 2};
        std::uint64_t id = 0;
        for (double a : vals)
            for (double b : vals) written.push_back(Particle(id++, a, b, -a, b, a, b));
        written[3].Drift(11, -1);
        written.saveAsTxt(textFile);
        Particles read;
        read.loadFromTxt(textFile);
        assert(read.size() == written.size());
        for (size_t q = 0; q < read.size(); ++q) {
            const auto &x = read[q].data(), &y = written[q].data();
            assert(x.id == y.id);
            assert(std::memcmp(&x.k, &y.k, 6 * sizeof(double)) == 0);
            assert(x.driftObjectType == y.driftObjectType && x.driftSide == y.driftSide);
        }
        // the header line is a tab and the count (Particles.cpp:28-31)
        std::ifstream in(textFile);
        std::string first;
        std::getline(in, first);
        assert(first == "\t" + std::to_string(written.size()));
    }
    // Fortran WaComM records: "i j k health age", id = line number, time = -1
    {
        std::ofstream out(textFile);
        out << "       2\n  10.5 20.25 -1.5 0.5 3600\n  11 21 -2 1 0\n";
        out.close();
        Particles read;
        read.loadFromTxt(textFile);
        assert(read.size() == 2);
        assert(read[0].Id() == 1 && read[0].I() == 10.5 && read[0].J() == 20.25 && read[0].K() == -1.5);
        assert(read[0].Health() == 0.5 && read[0].Age() == 3600 && read[0].Time() == -1);
        assert(read[1].Id() == 2 && read[1].DriftObject() == 0 && read[1].Side() == 0);
    }
    // a record of 6 tokens is rejected (Particles.cpp:354-356)
    {
        std::ofstream out(textFile);
        out << "\t1\n7 1 2 3 4 5\n";
        out.close();
        bool rejected = false;
        try {
            Particles read;
            read.loadFromTxt(textFile);
        } catch (const std::runtime_error &) {
            rejected = true;
        }
        assert(rejected);
    }
    std::remove(textFile.c_str());
    std::cout << "particles_txt_test: PASS" << std::endl;
    return 0;
}
