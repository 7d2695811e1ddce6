// fluid_test -- runs a drop-impact simulation from a .cfg file, writing the
// output directory <name>.odr (the reference's fluid_test.cc, with the step
// on the GPU).  With -r it resumes from the last restart file there.
#include <sys/stat.h>
#include <sys/types.h>

#include <cstdio>
#include <cstring>

#include "fluid_2d.hh"

int main(int argc, char **argv) {
    if (argc != 2 && (argc != 3 || strcmp(argv[1], "-r") != 0)) {
        fputs("Syntax: ./fluid_test [-r] <input_file>\n\n"
              "The input file should have a '.cfg' suffix\n",
              stderr);
        return 1;
    }
    fluid_2d f2d(argv[argc - 1]);
    mkdir(f2d.filename.c_str(), S_IRWXU | S_IRWXG | S_IROTH | S_IXOTH);
    if (argc == 2) f2d.init_fields();
    else f2d.load_last_restart();
    f2d.solve();
    return 0;
}
