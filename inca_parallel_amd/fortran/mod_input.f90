!! Namelist input and grid set-up, with the behaviour of the reference's
!! src/mod_input_reader.f90: the blocks &files, &intracule_job and &cube_job
!! (each optional; a syntax error stops the run), the same defaults, the
!! calc_type dispatch and the default output names derived from the input
!! file name.  &intracule_job also accepts engine options of this
!! build: engine_precision ('fp64' or 'mixed'), mixed_tail,
!! engine_polynomials ('reference' fits or 'exact' expansion),
!! point_cutoff (log2 of the per-point cutoff, 0 = evaluate every term),
!! engine_grouping (.false.: evaluate record by record instead of by
!! shell-quartet groups) and learned_grid_weights (a .npy of the learned autogrid network: predicted
!! radii and per-centre orders for the automatic multicentre grid, checked
!! against the sum rule with the plain autogrid as fallback).
module input_reader
   use inputdat
   use intrainfo
   use quadratures
   use autogrid
   implicit none

   real(8), dimension(20) :: scan_start = 0.0d0, scan_end = 0.0d0, scan_step = 0.0d0
   integer, dimension(20) :: scan_nang = 590
   integer :: nquad_in = 1
   real(8), dimension(3,50) :: cent_in = 0.0d0
   integer, dimension(50) :: nradc_in = 50, nangc_in = 590
   real(8), dimension(50) :: sfalpha_in = 1.0d0, Ps_in = 1.0d0
   logical :: multicenter = .false., manual_grid = .false.
   character(len=16) :: engine_precision = 'fp64'
   character(len=16) :: engine_polynomials = 'reference'

   ! &cube_job fields (one-electron cubes, mod_cubefile.f90 of this host)
   integer :: mo = 0, npr = 0, cao = 0
   double precision :: center(3) = 0.d0, step(3) = 0.d0
   integer :: np(3) = 0
   ! the input file name without its extension (names of the one-electron cubes)
   character(len=256) :: job_stem = ''

contains

   subroutine read_all_input(filename)
      character(len=*), intent(in) :: filename
      integer :: i, ios, dot_idx
      character(len=256) :: errmsg
      namelist /files/ wfxfilename, fchkfilename, basfilename, logfilename, dm2name, outname, &
                       cubeintraname, r_plot_name
      namelist /intracule_job/ calc_type, c1calc, intracule_at_zero, intracule_two_points, &
                               thresh, nosym, nohydro, x_point1, y_point1, z_point1, &
                               x_point2, y_point2, z_point2, multicenter, manual_grid, nrad, &
                               nang, nquad_in, cent_in, nradc_in, nangc_in, sfalpha_in, Ps_in, &
                               nblock, scan_start, scan_end, scan_step, scan_nang, center_i, &
                               step_i, np_i, engine_precision, mixed_tail, &
                               engine_polynomials, point_cutoff, engine_grouping, &
                               learned_grid_weights
      namelist /cube_job/ cube, denscube, primcube, aocube, mocube, gradient, laplacian, npr, &
                          cao, mo, center, step, np

      thresh = 1.0d-8
      intracalc = .false.; c1calc = .false.; cube = .false.
      denscube = .false.; primcube = .false.; aocube = .false.; mocube = .false.
      gradient = .false.; laplacian = .false.
      radial_integral = .false.; radial_plot = .false.; cubeintra = .false.
      calc_type = ""
      intracule_at_zero = .false.; intracule_two_points = .false.
      definite = .false.; multicenter = .true.; manual_grid = .false.
      betaone = .false.; nohydro = .true.; nosym = .false.
      nrad = 50; nang = 590; nblock = 1
      wfxfilename = ""; fchkfilename = ""; basfilename = ""; logfilename = ""
      dm2name = ""; outname = ""; cubeintraname = ""; r_plot_name = ""

      open(unit=10, file=filename, status='old', iostat=ios)
      if (ios /= 0) then
         write(*,*) "ERROR: Could not open input file: ", trim(filename)
         stop
      end if
      read(10, nml=files, iostat=ios, iomsg=errmsg)
      if (ios > 0) call bad_block("&files")
      rewind(10)
      read(10, nml=intracule_job, iostat=ios, iomsg=errmsg)
      if (ios > 0) call bad_block("&intracule_job")
      rewind(10)
      read(10, nml=cube_job, iostat=ios, iomsg=errmsg)
      if (ios > 0) call bad_block("&cube_job")
      close(10)

      readwfx = len_trim(wfxfilename) > 0
      readfchk = len_trim(fchkfilename) > 0
      readbas = len_trim(basfilename) > 0
      readlog = len_trim(logfilename) > 0

      call to_lower(calc_type)
      select case (trim(calc_type))
      case ("")
      case ("radial_integral")
         radial_integral = .true.
      case ("radial_plot")
         radial_plot = .true.
      case ("cubeintra")
         cubeintra = .true.
      case default
         write(*,*) "ERROR: Unknown calc_type '", trim(calc_type), "'."
         write(*,*) "Valid values: radial_integral, radial_plot, cubeintra."
         stop 1
      end select
      intracalc = (len_trim(calc_type) > 0) .or. intracule_at_zero .or. intracule_two_points

      call to_lower(engine_precision)
      select case (trim(engine_precision))
      case ("fp64")
         mixed_precision = .false.
      case ("mixed")
         mixed_precision = .true.
      case default
         write(*,*) "ERROR: Unknown engine_precision '", trim(engine_precision), "'."
         write(*,*) "Valid values: fp64, mixed."
         stop 1
      end select

      call to_lower(engine_polynomials)
      select case (trim(engine_polynomials))
      case ("reference")
         exact_polynomials = .false.
      case ("exact")
         exact_polynomials = .true.
      case default
         write(*,*) "ERROR: Unknown engine_polynomials '", trim(engine_polynomials), "'."
         write(*,*) "Valid values: reference, exact."
         stop 1
      end select

      if (.not. (point_cutoff == 0.0d0 .or. (point_cutoff >= 64.0d0 .and. point_cutoff <= 1000.0d0))) then
         write(*,*) "ERROR: point_cutoff must be 0 (off) or between 64 and 1000."
         stop 1
      end if

      if (.not. readwfx .and. .not. readfchk) then
         write(*,*) "ERROR: No wavefunction file specified. Please provide a .wfx or .fchk file."
         stop
      end if

      dot_idx = index(trim(filename), '.', back=.true.)
      if (len_trim(outname) == 0) outname = stem(filename, dot_idx)//".out"
      if (len_trim(r_plot_name) == 0) r_plot_name = stem(filename, dot_idx)//".rad"
      if (len_trim(cubeintraname) == 0) cubeintraname = stem(filename, dot_idx)//"_vector.cube"
      job_stem = stem(filename, dot_idx)

      if (radial_plot) then
         if (nblock < 1 .or. nblock > 20) then
            write(*,*) "ERROR: nblock must be between 1 and 20."
            stop 1
         end if
         allocate(n_an_per_part(nblock), tart(2,nblock), stp(nblock))
         do i = 1, nblock
            tart(1,i) = scan_start(i)
            tart(2,i) = scan_end(i)
            stp(i) = scan_step(i)
            n_an_per_part(i) = scan_nang(i)
         end do
      end if

   contains

      function stem(name, dot) result(s)
         character(len=*), intent(in) :: name
         integer, intent(in) :: dot
         character(len=:), allocatable :: s
         if (dot > 0) then
            s = name(1:dot-1)
         else
            s = trim(name)
         end if
      end function stem

      subroutine to_lower(str)
         character(len=*), intent(inout) :: str
         integer :: k
         do k = 1, len(str)
            if (str(k:k) >= 'A' .and. str(k:k) <= 'Z') str(k:k) = achar(iachar(str(k:k)) + 32)
         end do
      end subroutine to_lower

      subroutine bad_block(blockname)
         character(len=*), intent(in) :: blockname
         write(*,*) "ERROR: Syntax error in ", blockname, " block of the input file."
         write(*,*) trim(errmsg)
         stop 1
      end subroutine bad_block

   end subroutine read_all_input

   !! Integration centres of a radial_integral job (src/mod_input_reader.f90:185-221).
   subroutine setup_grids()
      integer :: i
      if (.not. radial_integral) return
      if (multicenter) then
         if (manual_grid) then
            nquad = nquad_in
            if (nquad < 1 .or. nquad > 50) then
               write(*,*) "ERROR: nquad_in must be between 1 and 50."
               stop 1
            end if
            allocate(cent(3,nquad), nradc(nquad), nangc(nquad), sfalpha(nquad), Ps(nquad))
            do i = 1, nquad
               cent(:,i) = cent_in(:,i)
               nradc(i) = nradc_in(i)
               nangc(i) = nangc_in(i)
               sfalpha(i) = sfalpha_in(i)
               Ps(i) = Ps_in(i)
            end do
         else if (len_trim(learned_grid_weights) > 0) then
            call learned_centercalc()
         else
            call centercalc()
         end if
      else
         nquad = 1
         allocate(cent(3,1), nradc(1), nangc(1), sfalpha(1), Ps(1))
         cent(:,1) = 0.0d0
         nradc(1) = nrad
         nangc(1) = nang
         sfalpha(1) = sfalpha_in(1)
         Ps(1) = 1.0d0
      end if
   end subroutine setup_grids

end module input_reader
