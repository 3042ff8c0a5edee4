// The command-line front end: every command is femm_cli_main of the library.

#include "femm_wire.h"

int main(int argc, char **argv)
{
  return femm_cli_main(argc, argv);
}
