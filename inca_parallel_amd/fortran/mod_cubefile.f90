!! One-electron cube files of &cube_job (behaviour of the reference's
!! src/mod_cubefile.f90:16-128 with the functions of src/mod_functions.f90),
!! evaluated on the GPU by inca_onee on the cube grid of gridpoints3:
!!   primcube  -> primitive npr (Prim), file <input>_prim.cube;
!!   mocube    -> orbital mo (MoOr; MO_a, MO_b and their sum for an
!!                unrestricted wavefunction), <input>_mo.cube;
!!   denscube  -> the density (Density; Dens_a, Dens_b and their sum),
!!                <input>_dens.cube;
!!   laplacian -> the reference's Lapl, <input>_lapl.cube.
!! Unrestricted runs write three files prefixed alpha, beta_ and all__ as the
!! reference does.  The reference never sets the names of these files
!! (nameprim, namemo, namedens, namelap are blank); this host names them as
!! above.  Atomic-orbital cubes (aocube) need the .log reader, which is not
!! part of this build.  Orbital coefficients come from the .wfx file
!! (wfxinfo T).
module cubefile
   use iso_c_binding
   use geninfo
   use wfxinfo
   use inputdat
   use input_reader, only: center, step, np, mo, npr, job_stem
   use inca_hip
   implicit none
   private
   public :: onee_cubes
contains
   subroutine onee_cubes()
      integer(c_int64_t) :: ng
      call inca_open()
      call upload_primitives()
      call inca_check(inca_grid_cube(inca_ctx, center, step, int(np, c_int32_t), ng), &
                      "inca_grid_cube")
      if (primcube) call write_one(1, trim(job_stem)//"_prim.cube", "")
      if (aocube) write(*,*) "NOTE: atomic-orbital cubes need the .log reader, " // &
                             "which is not part of this build."
      if (mocube) call spin_files(3, trim(job_stem)//"_mo.cube")
      if (denscube) call spin_files(5, trim(job_stem)//"_dens.cube")
      if (laplacian) call write_one(6, trim(job_stem)//"_lapl.cube", "")
   end subroutine onee_cubes

   subroutine upload_primitives()
      integer(c_int32_t), allocatable, target :: tmn32(:)
      allocate(tmn32(3*nprim))
      tmn32 = reshape(int(TMN(1:nprim,1:3), c_int32_t), (/3*nprim/))
      if (.not. allocated(Xn)) then
         allocate(Xn(nprim), Yn(nprim), Zn(nprim))
         Xn = cartes(Ra(1:nprim),1); Yn = cartes(Ra(1:nprim),2); Zn = cartes(Ra(1:nprim),3)
      end if
      call inca_check(inca_set_primitives(inca_ctx, int(nprim, c_int32_t), Alpha, tmn32, Xn, &
                      Yn, Zn, 0_c_int32_t), "inca_set_primitives")
   end subroutine upload_primitives

   !! Unrestricted orbital and density cubes come as alpha, beta and total
   !! files (mod_cubefile.f90:23-45); the rest as one file.
   subroutine spin_files(kind, name)
      integer, intent(in) :: kind
      character(len=*), intent(in) :: name
      if (uhf) then
         call write_one(kind, "alpha"//name, "alpha")
         call write_one(kind, "beta_"//name, "beta")
         call write_one(kind, "all__"//name, "all")
      else
         call write_one(kind, name, "")
      end if
   end subroutine spin_files

   subroutine write_one(kind, name, part)
      integer, intent(in) :: kind
      character(len=*), intent(in) :: name, part
      double precision, allocatable, target :: vals(:), tt(:,:), w(:), v2(:)
      integer :: i, nmo, m0, idx
      integer(c_int64_t) :: ng
      ng = int(np(1), c_int64_t)*np(2)*np(3)
      allocate(vals(max(ng, 1_c_int64_t)))
      if (kind == 1) then
         if (npr < 1 .or. npr > nprim) call stop_cube("npr outside 1..nprim")
         call inca_check(inca_onee(inca_ctx, 1_c_int32_t, int(npr, c_int32_t), 0_c_int32_t, &
                         c_null_ptr, c_null_ptr, c_loc(vals)), "inca_onee")
      else
         if (.not. allocated(T)) call stop_cube("orbital cubes need the MO coefficients of a .wfx file")
         ! the orbitals of this cube: all of them (closed shell), the alpha
         ! ones (rows 1..nalfae of T) or the beta ones (the next nbetae)
         if (uhf .and. part == "beta") then
            m0 = nalfae; nmo = nbetae
         else if (uhf .and. part == "alpha") then
            m0 = 0; nmo = nalfae
         else if (uhf) then
            m0 = 0; nmo = nalfae + nbetae
         else
            m0 = 0; nmo = noccmo
         end if
         allocate(tt(nmo, nprim), w(nmo))
         tt = T(m0+1:m0+nmo, 1:nprim)
         w = merge(1.d0, 2.d0, uhf)
         if (kind == 3) then
            if (mo < 1) call stop_cube("mo must be >= 1")
            if (uhf .and. part == "all") then
               ! MO_a(mo) + MO_b(mo)
               if (mo > nalfae .or. mo > nbetae) call stop_cube("mo outside the occupied orbitals")
               allocate(v2(max(ng, 1_c_int64_t)))
               call inca_check(inca_onee(inca_ctx, 3_c_int32_t, int(mo, c_int32_t), &
                               int(nmo, c_int32_t), c_loc(tt), c_loc(w), c_loc(vals)), "inca_onee")
               call inca_check(inca_onee(inca_ctx, 3_c_int32_t, int(nalfae + mo, c_int32_t), &
                               int(nmo, c_int32_t), c_loc(tt), c_loc(w), c_loc(v2)), "inca_onee")
               vals = vals + v2
            else
               idx = mo
               if (idx > nmo) call stop_cube("mo outside the occupied orbitals")
               call inca_check(inca_onee(inca_ctx, 3_c_int32_t, int(idx, c_int32_t), &
                               int(nmo, c_int32_t), c_loc(tt), c_loc(w), c_loc(vals)), "inca_onee")
            end if
         else
            call inca_check(inca_onee(inca_ctx, int(kind, c_int32_t), 0_c_int32_t, &
                            int(nmo, c_int32_t), c_loc(tt), c_loc(w), c_loc(vals)), "inca_onee")
         end if
      end if
      open(unit=2, file=name, status='replace')
      write(2,*) "CUBE FILE"
      write(2,*) "OUTER LOOP:X, MIDDLE LOOP:Y, INNER LOOP:Z"
      write(2,*) natoms, center(1), center(2), center(3)
      write(2,*) np(1), step(1), 0.d0, 0.d0
      write(2,*) np(2), 0.d0, step(2), 0.d0
      write(2,*) np(3), 0.d0, 0.d0, step(3)
      do i = 1, natoms
         write(2,*) an(i), chrg(i), cartes(i,1), cartes(i,2), cartes(i,3)
      end do
      do i = 1, int(ng)
         write(2,'(6(E16.6E3))') vals(i)
      end do
      close(2)
      write(*,*) "Cube file written: ", trim(name)
   end subroutine write_one

   subroutine stop_cube(msg)
      character(len=*), intent(in) :: msg
      write(*,*) "ERROR (&cube_job): ", msg
      stop 1
   end subroutine stop_cube
end module cubefile
