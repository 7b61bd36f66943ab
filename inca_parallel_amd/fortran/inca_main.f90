!! INCA driver (behaviour of the reference's src/inca_main.f90): read the
!! namelist input named on the command line, the wavefunction and basis
!! files, set up the integration centres and run calculate_intracule.
program inca
   use inputdat
   use intrainfo
   use input_reader
   use read_files
   use intracule
   use cubefile, only: onee_cubes
   use inca_hip, only: inca_close
   implicit none
   character(len=256) :: input_file_name

   if (command_argument_count() == 0) then
      write(*,*) "====================================================="
      write(*,*) " INCA: Intracule Calculator (MI355X engine)"
      write(*,*) "====================================================="
      write(*,*) " ERROR: No input file provided."
      write(*,*) " USAGE: inca <input_file>"
      write(*,*) "====================================================="
      stop
   end if
   call get_command_argument(1, input_file_name)
   call read_all_input(trim(input_file_name))

   if (readwfx) call filewfx(wfxfilename)
   if (readfchk) call filefchk(fchkfilename)
   if (readbas) call filebas(basfilename)
   if (readlog) write(*,*) "NOTE: .log input is not used by this build (intracules only)."
   if (cube) call onee_cubes()

   if (intracalc) then
      if (.not. readbas) then
         write(*,*) "CAUTION: Primitive info comes from .wfx or .fchk file"
         write(*,*) " i primtive may not coincide with that of the .dm2p file"
      end if
      call setup_grids()
      call calculate_intracule(.true.)
   end if
   if (c1calc) write(*,*) "NOTE: the exchange-hole model (c1calc) is not part of this build."
   call inca_close()
end program inca
